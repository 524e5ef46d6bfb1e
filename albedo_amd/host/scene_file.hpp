// Scene text format, OBJ meshes and PPM output for the C++ drop-in.
//
// Restates, by behaviour:
//   readScene, SceneLimits, SceneParseError, SceneDescription
//                          scene/scene_file.hpp:17-146, scene/scene_file.cpp:1-1114
//   readObj, ObjLimits, ObjParseError   geometry/obj.hpp, geometry/obj.cpp
//   Tokens, parseReal, parseInteger     io/tokens.cpp:10-50
//   encodeSrgb, toByte, writePpm        color/srgb.cpp:18-36, image/ppm.cpp:22-85
// Error messages keep the reference's wording ("scene line N: ...",
// "OBJ line N: ..."), numbers printed in shortest round-trip form.
#pragma once

#include "albedo_dropin.hpp"

#include <charconv>
#include <filesystem>
#include <fstream>
#include <istream>
#include <map>
#include <sstream>
#include <string_view>

namespace albedo {

// ----------------------------------------------------------------- tokens

class Tokens {
public:
    explicit Tokens(std::string_view text) : rest_(text) {}
    std::optional<std::string_view> next()
    {
        constexpr std::string_view ws{" \t\r\v\f"};
        const std::size_t start = rest_.find_first_not_of(ws);
        if (start == std::string_view::npos) {
            rest_ = {};
            return std::nullopt;
        }
        rest_.remove_prefix(start);
        const std::size_t end = std::min(rest_.find_first_of(ws), rest_.size());
        const std::string_view tok = rest_.substr(0, end);
        rest_.remove_prefix(end);
        return tok;
    }

private:
    std::string_view rest_;
};

inline std::optional<Real> parseReal(std::string_view token)
{
    if (!token.empty() && token.front() == '+') {
        token.remove_prefix(1);
        if (!token.empty() && (token.front() == '+' || token.front() == '-')) return std::nullopt;
    }
    Real value{};
    const auto [end, ec] = std::from_chars(token.data(), token.data() + token.size(), value);
    if (ec != std::errc{} || end != token.data() + token.size() || !std::isfinite(value)) return std::nullopt;
    return value;
}

inline std::optional<long long> parseInteger(std::string_view token)
{
    long long value{};
    const auto [end, ec] = std::from_chars(token.data(), token.data() + token.size(), value);
    if (ec != std::errc{} || end != token.data() + token.size() || token.empty()) return std::nullopt;
    return value;
}

namespace detail {

// A double in shortest round-trip form, as std::format("{}") prints it.
inline std::string num(Real v)
{
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, v);
    return std::string(buf, r.ptr);
}

// `token` in quotes, cut to 32 bytes, non-printable bytes as \xNN
// (reference scene_file.cpp quoted()).
inline std::string quoted(std::string_view token)
{
    std::string out{"'"};
    for (const char c : token.substr(0, 32)) {
        const auto b = static_cast<unsigned char>(c);
        if (b >= 0x20 && b <= 0x7E) {
            out += c;
        } else {
            char hex[8];
            std::snprintf(hex, sizeof hex, "\\x%02X", b);
            out += hex;
        }
    }
    out += token.size() > 32 ? "'..." : "'";
    return out;
}

inline bool isIdentifier(std::string_view s)
{
    auto letter = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); };
    auto digit = [](char c) { return c >= '0' && c <= '9'; };
    if (s.empty() || (!letter(s.front()) && s.front() != '_')) return false;
    for (char c : s)
        if (!(letter(c) || digit(c) || c == '_' || c == '-')) return false;
    return true;
}

// Lines of at most `limit` bytes (without the terminator), numbered from 1.
class LineReader {
public:
    LineReader(std::istream& in, std::size_t limit) : in_(in), limit_(limit) {}
    bool next(std::string& line)
    {
        line.clear();
        int c;
        bool any = false;
        while ((c = in_.get()) != EOF) {
            any = true;
            if (c == '\n') break;
            if (line.size() >= limit_) {
                ++number_;
                throw std::length_error(std::to_string(number_));
            }
            line.push_back(char(c));
        }
        if (!any) {
            if (in_.bad()) throw std::runtime_error("cannot read input");
            return false;
        }
        ++number_;
        return true;
    }
    std::size_t number() const { return number_; }

private:
    std::istream& in_;
    std::size_t limit_;
    std::size_t number_{0};
};

} // namespace detail

// ------------------------------------------------------------------- OBJ

struct ObjLimits {
    std::size_t max_line_length{65536};
    std::size_t max_vertices{std::size_t{1} << 24U};
    std::size_t max_texture_coordinates{std::size_t{1} << 24U};
    std::size_t max_triangles{std::size_t{1} << 24U};
};

class ObjParseError : public std::runtime_error {
public:
    ObjParseError(std::size_t line, std::string_view problem)
        : std::runtime_error("OBJ line " + std::to_string(line) + ": " + std::string(problem)), line_(line)
    {
    }
    std::size_t line() const noexcept { return line_; }

private:
    std::size_t line_;
};

// reference geometry/obj.cpp:41-238: v, vt, f (polygons as fans, negative
// indices from the end, optional texture indices); other statements ignored.
inline TriangleMesh readObj(std::istream& input, const ObjLimits& limits = {})
{
    detail::LineReader lines(input, limits.max_line_length);
    std::vector<Point3r> pos;
    std::vector<Point2r> uv;
    std::vector<TriangleMesh::Face> faces, uv_faces;
    bool every_textured = true;
    std::string line;
    auto fail = [&](const std::string& problem) -> void { throw ObjParseError(lines.number(), problem); };
    auto coord = [&](std::string_view t) {
        const auto v = parseReal(t);
        if (!v) fail("a coordinate is not a finite number");
        return *v;
    };
    auto resolved = [&](long long index, const char* kind, std::size_t count) -> std::size_t {
        if (index > 0) {
            if (static_cast<unsigned long long>(index) > count)
                fail(std::string("a face refers to a ") + kind + " not yet defined");
            return std::size_t(index - 1);
        }
        if (index < 0) {
            const unsigned long long back = static_cast<unsigned long long>(-(index + 1)) + 1;
            if (back > count) fail(std::string("a face refers to a ") + kind + " before the first one");
            return count - std::size_t(back);
        }
        fail("a face uses index 0; OBJ indices start at 1");
        return 0;
    };
    try {
        while (lines.next(line)) {
            Tokens tok(std::string_view(line).substr(0, line.find('#')));
            const auto kw = tok.next();
            if (kw == "v") {
                if (pos.size() >= limits.max_vertices)
                    fail("more than " + std::to_string(limits.max_vertices) + " vertices");
                Real xyz[3];
                for (Real& c : xyz) {
                    const auto t = tok.next();
                    if (!t) fail("a vertex needs three coordinates");
                    c = coord(*t);
                }
                while (const auto extra = tok.next()) (void)coord(*extra);
                pos.push_back({xyz[0], xyz[1], xyz[2]});
            } else if (kw == "vt") {
                if (uv.size() >= limits.max_texture_coordinates)
                    fail("more than " + std::to_string(limits.max_texture_coordinates) + " texture coordinates");
                const auto u = tok.next();
                if (!u) fail("a texture coordinate needs at least u");
                Point2r p{coord(*u), 0.0};
                if (const auto v = tok.next()) p.y = coord(*v);
                while (const auto extra = tok.next()) (void)coord(*extra);
                uv.push_back(p);
            } else if (kw == "f") {
                struct Corner {
                    std::size_t p;
                    std::optional<std::size_t> t;
                };
                std::vector<Corner> corners;
                while (const auto t = tok.next()) {
                    const std::size_t s1 = t->find('/');
                    const auto pi = parseInteger(t->substr(0, s1));
                    if (!pi) fail("a face corner is not a valid index");
                    Corner c{resolved(*pi, "vertex", pos.size()), std::nullopt};
                    if (s1 != std::string_view::npos) {
                        const std::string_view rest = t->substr(s1 + 1);
                        const std::size_t s2 = rest.find('/');
                        const std::string_view tp = rest.substr(0, s2);
                        if (!tp.empty()) {
                            const auto ti = parseInteger(tp);
                            if (!ti) fail("a face corner is not a valid index");
                            c.t = resolved(*ti, "texture coordinate", uv.size());
                        }
                        if (s2 != std::string_view::npos) {
                            const std::string_view np = rest.substr(s2 + 1);
                            if (!np.empty() && !parseInteger(np)) fail("a face corner is not a valid index");
                        }
                    }
                    corners.push_back(c);
                }
                if (corners.size() < 3) fail("a face needs at least three corners");
                if (corners.size() - 2 > limits.max_triangles - faces.size())
                    fail("more than " + std::to_string(limits.max_triangles) + " triangles");
                for (std::size_t k = 1; k + 1 < corners.size(); ++k) {
                    const Corner &a = corners[0], &b = corners[k], &c = corners[k + 1];
                    faces.push_back({a.p, b.p, c.p});
                    if (!a.t || !b.t || !c.t) {
                        every_textured = false;
                        uv_faces.clear();
                    } else if (every_textured) {
                        uv_faces.push_back({*a.t, *b.t, *c.t});
                    }
                }
            }
        }
    } catch (const std::length_error& e) {
        throw ObjParseError(std::stoull(e.what()),
                            "line longer than " + std::to_string(limits.max_line_length) + " bytes");
    }
    return TriangleMesh(std::move(pos), std::move(faces), std::move(uv), std::move(uv_faces));
}

// ----------------------------------------------------------------- scenes

struct SceneLimits {
    std::size_t max_line_length{4096};
    std::size_t max_objects{4096};
    std::size_t max_materials{4096};
    std::size_t max_meshes{16};
    ObjLimits mesh_limits{};
    std::size_t max_instance_operations{16};
    std::size_t max_ground_points{std::size_t{1} << 26U};
};

class SceneParseError : public std::runtime_error {
public:
    SceneParseError(std::size_t line, std::string_view problem)
        : std::runtime_error(line == 0 ? "scene: " + std::string(problem)
                                       : "scene line " + std::to_string(line) + ": " + std::string(problem)),
          line_(line)
    {
    }
    std::size_t line() const noexcept { return line_; }

private:
    std::size_t line_;
};

struct SceneDescription {
    Scene scene;
    CameraSettings camera;
};

namespace detail {

constexpr Real kMaxReflectance = 1, kMaxEmission = 1e6, kMaxRoughness = 1, kMinIor = 1e-2, kMaxIor = 1e2,
               kMinEta = 1e-3, kMaxEta = 1e3, kMaxK = 1e3, kMaxVeins = 1e2;

class Statement {
public:
    Statement(std::string_view text, std::size_t line) : tokens_(text), line_(line) {}
    [[noreturn]] void fail(const std::string& problem) const { throw SceneParseError(line_, problem); }
    std::optional<std::string_view> next() { return tokens_.next(); }
    std::optional<std::string_view> peek() const
    {
        Tokens rest = tokens_;
        return rest.next();
    }
    std::string_view word(const std::string& what)
    {
        const auto t = next();
        if (!t) fail("expected " + what);
        return *t;
    }
    Real number(const std::string& what)
    {
        const std::string_view t = word(what);
        const auto v = parseReal(t);
        if (!v) fail(what + " must be a finite number, got " + quoted(t));
        return *v;
    }
    Real inRange(const std::string& what, Real lo, Real hi)
    {
        const Real v = number(what);
        if (v < lo || v > hi) fail(what + " must be between " + num(lo) + " and " + num(hi) + ", got " + num(v));
        return v;
    }
    Real positive(const std::string& what)
    {
        const Real v = number(what);
        if (!(v > 0)) fail(what + " must be positive, got " + num(v));
        return v;
    }
    Point3r point(const std::string& what)
    {
        const Real x = number("the x coordinate of " + what);
        const Real y = number("the y coordinate of " + what);
        return {x, y, number("the z coordinate of " + what)};
    }
    Vector3r vector(const std::string& what)
    {
        const Point3r p = point(what);
        return {p.x, p.y, p.z};
    }
    Rgb color(const std::string& what, Real lo, Real hi)
    {
        const Real r = inRange("the red channel of " + what, lo, hi);
        const Real g = inRange("the green channel of " + what, lo, hi);
        return {r, g, inRange("the blue channel of " + what, lo, hi)};
    }
    Rgb reflectance(const std::string& what) { return color(what, 0, kMaxReflectance); }
    Rgb emission(const std::string& what) { return color(what, 0, kMaxEmission); }
    std::string_view name(const std::string& what)
    {
        const std::string_view t = word(what);
        if (!isIdentifier(t)) fail(quoted(t) + " is not a valid name");
        return t;
    }
    template <class Read>
    void readFields(Read read)
    {
        std::vector<std::string_view> seen;
        while (const auto f = next()) {
            if (std::find(seen.begin(), seen.end(), *f) != seen.end()) fail(quoted(*f) + " is given twice");
            seen.push_back(*f);
            if (!read(*f)) fail("unknown field " + quoted(*f));
        }
    }
    template <class T>
    T required(const std::optional<T>& v, const char* what) const
    {
        if (!v) fail(std::string("missing '") + what + "'");
        return *v;
    }
    void expectEnd()
    {
        if (const auto extra = next()) fail("unexpected " + quoted(*extra));
    }

private:
    Tokens tokens_;
    std::size_t line_;
};

inline bool isInstanceKeyword(std::string_view w)
{
    for (const char* k : {"material", "translate", "scale", "rotate", "fit", "ground"})
        if (w == k) return true;
    return false;
}

struct LoadedMesh {
    std::shared_ptr<const MeshGeometry> geometry;
};

// The reader of scene/scene_file.cpp:443-1050.
class SceneReader {
public:
    SceneReader(const SceneLimits& limits, std::optional<std::filesystem::path> base)
        : limits_(limits), base_(std::move(base)), mesh_budget_(limits.mesh_limits),
          ground_left_(limits.max_ground_points)
    {
    }

    void readLine(std::string_view line, std::size_t number)
    {
        Statement st(line.substr(0, line.find('#')), number);
        const auto kw = st.next();
        if (!kw) return;
        if (*kw == "camera") readCamera(st);
        else if (*kw == "sky") readSky(st);
        else if (*kw == "material") readMaterial(st);
        else if (*kw == "sphere") readSphere(st);
        else if (*kw == "point_light") readPointLight(st);
        else if (*kw == "area_light") readAreaLight(st);
        else if (*kw == "mesh") readMesh(st);
        else if (*kw == "instance") readInstance(st);
        else st.fail("unknown statement " + quoted(*kw));
    }

    SceneDescription finish()
    {
        if (!camera_) throw SceneParseError(0, "the scene has no camera");
        return {std::move(scene_), *camera_};
    }

private:
    void readCamera(Statement& st)
    {
        if (camera_) st.fail("the scene already has a camera");
        std::optional<Point3r> from, at;
        CameraSettings s{};
        st.readFields([&](std::string_view f) {
            if (f == "from") from = st.point("the camera position");
            else if (f == "at") at = st.point("the point the camera looks at");
            else if (f == "up") s.up = st.vector("the camera's up direction");
            else if (f == "fov") s.vertical_fov_degrees = st.number("the field of view");
            else return false;
            return true;
        });
        s.look_from = st.required(from, "from");
        s.look_at = st.required(at, "at");
        checkCamera(st, s);
        camera_ = s;
    }

    // reference camera/perspective_camera.cpp:24-110
    static void checkCamera(const Statement& st, const CameraSettings& s)
    {
        const Real fov = s.vertical_fov_degrees;
        if (std::isnan(fov) || fov <= 0 || fov >= 180)
            st.fail("vertical field of view must be in (0, 180) degrees, got " + num(fov));
        const Real b[3] = {s.look_from.x - s.look_at.x, s.look_from.y - s.look_at.y, s.look_from.z - s.look_at.z};
        const Real bl2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        if (!std::isnormal(bl2))
            st.fail("distance from camera position to target must be non-zero, finite and within roughly "
                    "[1e-154, 1e154]");
        const Real ul2 = s.up.x * s.up.x + s.up.y * s.up.y + s.up.z * s.up.z;
        if (!std::isnormal(ul2))
            st.fail("length of the camera up direction must be non-zero, finite and within roughly [1e-154, 1e154]");
        const Real bl = std::sqrt(bl2), ul = std::sqrt(ul2);
        const Real w[3] = {b[0] / bl, b[1] / bl, b[2] / bl}, u[3] = {s.up.x / ul, s.up.y / ul, s.up.z / ul};
        const Real r[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < 1e-12)
            st.fail("camera up direction must not be parallel to the view");
    }

    void readSky(Statement& st)
    {
        if (sky_given_) st.fail("the scene already has a sky");
        const auto first = st.peek();
        if (first && (*first == "nadir" || *first == "zenith")) {
            std::optional<Rgb> nadir, zenith;
            st.readFields([&](std::string_view f) {
                if (f == "nadir") nadir = st.emission("the radiance from below");
                else if (f == "zenith") zenith = st.emission("the radiance from above");
                else return false;
                return true;
            });
            scene_.setSky(SkyLight{st.required(nadir, "nadir"), st.required(zenith, "zenith")});
        } else {
            scene_.setSky(SkyLight::uniform(st.emission("the sky's radiance")));
            st.expectEnd();
        }
        sky_given_ = true;
    }

    void readMaterial(Statement& st)
    {
        if (materials_.size() >= limits_.max_materials)
            st.fail("more than " + std::to_string(limits_.max_materials) + " materials");
        const std::string_view name = st.name("a material name");
        if (materials_.count(std::string(name))) st.fail("material " + quoted(name) + " is already defined");
        const std::string_view type = st.word("a material type");
        Material m = type == "mix" ? readMix(st) : ofType(st, type);
        materials_.emplace(std::string(name), m);
    }

    static Material ofType(Statement& st, std::string_view type)
    {
        if (type == "diffuse") {
            std::optional<Rgb> color;
            st.readFields([&](std::string_view f) {
                if (f != "color") return false;
                color = st.reflectance("the color");
                return true;
            });
            return DiffuseMaterial{ConstantTexture{st.required(color, "color")}};
        }
        if (type == "checker") {
            std::optional<Rgb> even, odd;
            std::optional<Real> freq;
            st.readFields([&](std::string_view f) {
                if (f == "even") even = st.reflectance("the even color");
                else if (f == "odd") odd = st.reflectance("the odd color");
                else if (f == "frequency") freq = st.positive("the checker frequency");
                else return false;
                return true;
            });
            return DiffuseMaterial{CheckerTexture{st.required(even, "even"), st.required(odd, "odd"),
                                                  st.required(freq, "frequency")}};
        }
        if (type == "perlin") {
            std::optional<Rgb> low, high;
            std::optional<Real> scale;
            Real veins = 0;
            st.readFields([&](std::string_view f) {
                if (f == "low") low = st.reflectance("the low color");
                else if (f == "high") high = st.reflectance("the high color");
                else if (f == "scale") scale = st.positive("the noise scale");
                else if (f == "veins") veins = st.inRange("the veins", 0, kMaxVeins);
                else return false;
                return true;
            });
            return DiffuseMaterial{PerlinTexture{st.required(low, "low"), st.required(high, "high"),
                                                 st.required(scale, "scale"), veins}};
        }
        if (type == "conductor") {
            std::optional<Rgb> eta, k;
            Real rough = 0;
            st.readFields([&](std::string_view f) {
                if (f == "eta") eta = st.color("the refractive index", kMinEta, kMaxEta);
                else if (f == "k") k = st.color("the absorption", 0, kMaxK);
                else if (f == "roughness") rough = st.inRange("the roughness", 0, kMaxRoughness);
                else return false;
                return true;
            });
            const ConductorMaterial m{st.required(eta, "eta"), st.required(k, "k"), rough};
            const Real e[3] = {m.eta.r, m.eta.g, m.eta.b}, kk[3] = {m.k.r, m.k.g, m.k.b};
            for (int c = 0; c < 3; ++c)
                if (e[c] == 1 && kk[c] == 0) st.fail("a conductor channel cannot have eta 1 and k 0");
            return m;
        }
        if (type == "dielectric") {
            std::optional<Real> ior;
            Real rough = 0;
            st.readFields([&](std::string_view f) {
                if (f == "ior") ior = st.inRange("the index of refraction", kMinIor, kMaxIor);
                else if (f == "roughness") rough = st.inRange("the roughness", 0, kMaxRoughness);
                else return false;
                return true;
            });
            return DielectricMaterial{st.required(ior, "ior"), rough};
        }
        st.fail("unknown material type " + quoted(type));
    }

    Material readMix(Statement& st) const
    {
        std::optional<BasicMaterial> first, second;
        std::optional<Real> scale;
        Real veins = 0, threshold = 0.5;
        st.readFields([&](std::string_view f) {
            if (f == "first") first = basicNamed(st);
            else if (f == "second") second = basicNamed(st);
            else if (f == "scale") scale = st.positive("the noise scale");
            else if (f == "veins") veins = st.inRange("the veins", 0, kMaxVeins);
            else if (f == "threshold") threshold = st.inRange("the threshold", 0, 1);
            else return false;
            return true;
        });
        return MixMaterial{st.required(first, "first"), st.required(second, "second"),
                           PerlinPattern{st.required(scale, "scale"), veins}, threshold};
    }

    BasicMaterial basicNamed(Statement& st) const
    {
        const std::string_view name = st.peek().value_or("");
        const Material& m = materialNamed(st);
        if (std::holds_alternative<MixMaterial>(m))
            st.fail("material " + quoted(name) + " is a mix, and a mix cannot hold another mix");
        if (const auto* d = std::get_if<DiffuseMaterial>(&m)) return *d;
        if (const auto* c = std::get_if<ConductorMaterial>(&m)) return *c;
        return std::get<DielectricMaterial>(m);
    }

    const Material& materialNamed(Statement& st) const
    {
        const std::string_view name = st.name("a material name");
        const auto it = materials_.find(std::string(name));
        if (it == materials_.end()) st.fail("material " + quoted(name) + " is not defined");
        return it->second;
    }

    void countObject(const Statement& st)
    {
        if (objects_ >= limits_.max_objects) st.fail("more than " + std::to_string(limits_.max_objects) + " objects");
        ++objects_;
    }

    static Sphere sphereAt(const Statement& st, Point3r c, Real r)
    {
        try {
            return Sphere{c, r};
        } catch (const std::invalid_argument& e) {
            st.fail(e.what());
        }
    }

    void readSphere(Statement& st)
    {
        countObject(st);
        const Point3r c = st.point("the sphere center");
        std::optional<Real> radius;
        const Material* m = nullptr;
        st.readFields([&](std::string_view f) {
            if (f == "radius") radius = st.number("a radius");
            else if (f == "material") m = &materialNamed(st);
            else return false;
            return true;
        });
        const Sphere s = sphereAt(st, c, st.required(radius, "radius"));
        if (!m) st.fail("missing 'material'");
        scene_.add(s, *m);
    }

    void readPointLight(Statement& st)
    {
        countObject(st);
        const Point3r p = st.point("the light position");
        std::optional<Rgb> intensity;
        st.readFields([&](std::string_view f) {
            if (f != "intensity") return false;
            intensity = st.emission("the intensity");
            return true;
        });
        scene_.addLight(PointLight{p, st.required(intensity, "intensity")});
    }

    void readAreaLight(Statement& st)
    {
        countObject(st);
        const Point3r c = st.point("the light center");
        std::optional<Real> radius;
        std::optional<Rgb> radiance;
        Material m{DiffuseMaterial{}};
        st.readFields([&](std::string_view f) {
            if (f == "radius") radius = st.number("a radius");
            else if (f == "radiance") radiance = st.emission("the radiance");
            else if (f == "material") m = materialNamed(st);
            else return false;
            return true;
        });
        const Sphere s = sphereAt(st, c, st.required(radius, "radius"));
        scene_.addAreaLight(s, st.required(radiance, "radiance"), m);
    }

    // Path rules of reference scene_file.cpp relativeMeshPath/meshFile.
    std::filesystem::path meshFile(const Statement& st, std::string_view text) const
    {
        for (char c : text) {
            const auto b = static_cast<unsigned char>(c);
            if (b < 0x20 || b == 0x7F) st.fail("mesh path " + quoted(text) + " must be UTF-8 without control characters");
        }
        const std::filesystem::path rel{std::string(text)};
        bool dotdot = false;
        for (const auto& part : rel)
            if (part == "..") dotdot = true;
        if (rel.is_absolute() || rel.has_root_name() || rel.has_root_directory() || dotdot)
            st.fail("mesh path " + quoted(text) + " must be relative to the scene and not use '..'");
        for (const auto& part : rel) {
            std::string base = part.string();
            base.erase(std::min(base.find_first_of(".:"), base.size()));
            while (!base.empty() && base.back() == ' ') base.pop_back();
            for (char& c : base)
                if (c >= 'a' && c <= 'z') c = char(c - 'a' + 'A');
            const bool plain = base == "CON" || base == "PRN" || base == "AUX" || base == "NUL";
            const bool numbered = (base.rfind("COM", 0) == 0 || base.rfind("LPT", 0) == 0) && base.size() == 4 &&
                                  base[3] >= '1' && base[3] <= '9';
            if (plain || numbered) st.fail("mesh path " + quoted(text) + " names a device");
        }
        std::error_code ec;
        const auto real_base = std::filesystem::canonical(*base_, ec);
        if (ec) st.fail("the scene's directory cannot be resolved");
        const auto real_file = std::filesystem::weakly_canonical(real_base / rel, ec);
        if (ec) st.fail("cannot resolve mesh path " + quoted(text));
        auto bi = real_base.begin();
        auto fi = real_file.begin();
        for (; bi != real_base.end() && fi != real_file.end() && *bi == *fi; ++bi, ++fi) {
        }
        if (bi != real_base.end()) st.fail("mesh path " + quoted(text) + " leads outside the scene's directory");
        const auto status = std::filesystem::status(real_file, ec);
        if (status.type() == std::filesystem::file_type::not_found)
            st.fail("mesh file " + quoted(text) + " does not exist");
        if (status.type() != std::filesystem::file_type::regular)
            st.fail("mesh path " + quoted(text) + " is not a regular file");
        return real_file;
    }

    void readMesh(Statement& st)
    {
        if (!base_) st.fail("meshes can only be read from a scene file, which they sit beside");
        if (meshes_.size() >= limits_.max_meshes) st.fail("more than " + std::to_string(limits_.max_meshes) + " meshes");
        const std::string_view name = st.name("a mesh name");
        if (meshes_.count(std::string(name))) st.fail("mesh " + quoted(name) + " is already defined");
        const std::string_view path = st.word("a mesh file");
        const bool smooth = st.peek() == std::optional<std::string_view>("smooth");
        if (smooth) (void)st.next();
        st.expectEnd();
        const auto file = meshFile(st, path);
        std::ifstream in(file, std::ios::binary);
        if (!in) st.fail("cannot read mesh file " + quoted(path));
        std::optional<TriangleMesh> mesh;
        try {
            mesh.emplace(readObj(in, mesh_budget_));
        } catch (const ObjParseError& e) {
            st.fail("mesh " + quoted(path) + ": " + e.what());
        } catch (const std::runtime_error&) {
            st.fail("cannot read mesh file " + quoted(path));
        } catch (const std::out_of_range& e) {
            st.fail("mesh " + quoted(path) + ": " + e.what());
        }
        std::shared_ptr<const MeshGeometry> geo;
        try {
            geo = std::make_shared<const MeshGeometry>(*mesh, smooth ? MeshShading::Smooth : MeshShading::Flat);
        } catch (const std::invalid_argument&) {
            st.fail("mesh " + quoted(path) + " has no triangles");
        }
        mesh_budget_.max_vertices -= mesh->positions().size();
        mesh_budget_.max_texture_coordinates -= mesh->textureCoordinates().size();
        mesh_budget_.max_triangles -= mesh->faces().size();
        meshes_.emplace(std::string(name), LoadedMesh{std::move(geo)});
    }

    Transform applied(Statement& st, std::string_view op, const LoadedMesh& mesh, const Transform& placement)
    {
        try {
            if (op == "translate") return Transform::translation(st.vector("the translation")) * placement;
            if (op == "scale") {
                const Real first = st.number("a scale factor");
                const auto next = st.peek();
                if (!next || isInstanceKeyword(*next)) return Transform::scaling({first, first, first}) * placement;
                const Real y = st.number("the y scale factor");
                return Transform::scaling({first, y, st.number("the z scale factor")}) * placement;
            }
            if (op == "rotate") {
                const Vector3r axis = st.vector("the rotation axis");
                return Transform::rotation(axis, st.number("the rotation angle")) * placement;
            }
            if (op == "fit") {
                const Real size = st.positive("the fitted size");
                Point3r lo, hi;
                placement.applyBox(mesh.geometry->boundsMin(), mesh.geometry->boundsMax(), lo, hi);
                if (!std::isfinite(lo.x) || !std::isfinite(lo.y) || !std::isfinite(lo.z) || !std::isfinite(hi.x) ||
                    !std::isfinite(hi.y) || !std::isfinite(hi.z))
                    st.fail("the instance is out of range before 'fit'");
                const Real ext = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
                const Real scale = size / ext;
                if (!std::isfinite(scale) || !(scale > 0)) st.fail("the instance is too large or too small to fit");
                const Vector3r centre{lo.x / 2 + hi.x / 2, lo.y / 2 + hi.y / 2, lo.z / 2 + hi.z / 2};
                return Transform::scaling({scale, scale, scale}) *
                       Transform::translation({-centre.x, -centre.y, -centre.z}) * placement;
            }
            if (op == "ground") {
                const auto& pts = mesh.geometry->usedPoints();
                if (pts.size() > ground_left_)
                    st.fail("'ground' would examine more than " + std::to_string(limits_.max_ground_points) +
                            " points in the scene");
                ground_left_ -= pts.size();
                std::optional<Real> lowest;
                for (const Point3r& p : pts) {
                    const Real h = placement.apply(p).y;
                    if (!std::isfinite(h)) st.fail("the instance is out of range before 'ground'");
                    lowest = lowest ? std::min(*lowest, h) : h;
                }
                return Transform::translation({0, -*lowest, 0}) * placement;
            }
        } catch (const std::invalid_argument& e) {
            st.fail(e.what());
        }
        st.fail("unknown instance operation " + quoted(op));
    }

    void readInstance(Statement& st)
    {
        countObject(st);
        const std::string_view mname = st.name("a mesh name");
        const auto it = meshes_.find(std::string(mname));
        if (it == meshes_.end()) st.fail("mesh " + quoted(mname) + " is not defined");
        const Material* m = nullptr;
        Transform placement;
        std::size_t ops = 0;
        while (const auto w = st.next()) {
            if (*w == "material") {
                if (m) st.fail("'material' is given twice");
                m = &materialNamed(st);
                continue;
            }
            if (++ops > limits_.max_instance_operations)
                st.fail("more than " + std::to_string(limits_.max_instance_operations) + " operations");
            placement = applied(st, *w, it->second, placement);
        }
        if (!m) st.fail("missing 'material'");
        Point3r lo, hi;
        placement.applyBox(it->second.geometry->boundsMin(), it->second.geometry->boundsMax(), lo, hi);
        if (!std::isfinite(lo.x) || !std::isfinite(lo.y) || !std::isfinite(lo.z) || !std::isfinite(hi.x) ||
            !std::isfinite(hi.y) || !std::isfinite(hi.z))
            st.fail("mesh instance transform takes the mesh out of range");
        scene_.add(MeshInstance{it->second.geometry, placement}, *m);
    }

    SceneLimits limits_;
    std::optional<std::filesystem::path> base_;
    ObjLimits mesh_budget_;
    std::size_t ground_left_;
    std::map<std::string, LoadedMesh> meshes_;
    std::map<std::string, Material> materials_;
    Scene scene_;
    std::optional<CameraSettings> camera_;
    bool sky_given_{false};
    std::size_t objects_{0};
};

inline SceneDescription readSceneFrom(std::istream& input, std::optional<std::filesystem::path> base,
                                      const SceneLimits& limits)
{
    if (base && base->empty()) base = ".";
    LineReader lines(input, limits.max_line_length);
    SceneReader reader(limits, std::move(base));
    std::string line;
    try {
        while (lines.next(line)) reader.readLine(line, lines.number());
    } catch (const std::length_error& e) {
        throw SceneParseError(std::stoull(e.what()),
                              "line longer than " + std::to_string(limits.max_line_length) + " bytes");
    }
    return reader.finish();
}

} // namespace detail

inline SceneDescription readScene(std::istream& input, const std::filesystem::path& base_directory,
                                  const SceneLimits& limits = {})
{
    return detail::readSceneFrom(input, base_directory, limits);
}

inline SceneDescription readScene(std::istream& input, const SceneLimits& limits = {})
{
    return detail::readSceneFrom(input, std::nullopt, limits);
}

inline SceneDescription readScene(const std::filesystem::path& path, const SceneLimits& limits = {})
{
    std::error_code ec;
    const auto type = std::filesystem::status(path, ec).type();
    if (!ec && type != std::filesystem::file_type::not_found && type != std::filesystem::file_type::regular)
        throw std::runtime_error("scene file '" + path.string() + "' is not a regular file");
    std::ifstream file(path, std::ios::binary);
    if (!file) throw std::runtime_error("cannot open scene file '" + path.string() + "'");
    return readScene(file, path.parent_path(), limits);
}

// ----------------------------------------------------------------- output

// sRGB transfer function (IEC 61966-2-1), reference color/srgb.cpp:18-36.
inline Real encodeSrgb(Real linear)
{
    if (linear <= 0.0031308) return 12.92 * linear;
    return 1.055 * std::pow(linear, 1 / 2.4) - 0.055;
}

inline std::uint8_t toByte(Real encoded)
{
    if (!(encoded > 0)) return 0;
    if (encoded >= 1) return 255;
    return static_cast<std::uint8_t>(std::lround(encoded * 255));
}

// Binary P6 of the per-pixel means (reference image/ppm.cpp:49-85).
inline void writePpm(std::ostream& out, const PixelBuffer& image)
{
    out << "P6\n" << image.width() << ' ' << image.height() << "\n255\n";
    std::vector<char> row(std::size_t(image.width()) * 3);
    for (int y = 0; y < image.height(); ++y) {
        for (int x = 0; x < image.width(); ++x) {
            const Rgb c = image.average({x, y});
            row[3 * x] = char(toByte(encodeSrgb(c.r)));
            row[3 * x + 1] = char(toByte(encodeSrgb(c.g)));
            row[3 * x + 2] = char(toByte(encodeSrgb(c.b)));
        }
        out.write(row.data(), std::streamsize(row.size()));
    }
    if (!out) throw std::runtime_error("cannot write the image");
}

inline void writePpm(const std::filesystem::path& path, const PixelBuffer& image)
{
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot open '" + path.string() + "' for writing");
    writePpm(out, image);
}

} // namespace albedo
