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

#include <algorithm>
#include <array>
#include <charconv>
#include <filesystem>
#include <fstream>
#include <istream>
#include <limits>
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

} // namespace detail

// A line over a reader's length limit (reference io/line_reader.hpp:14-31).
class LineTooLongError : public std::runtime_error {
public:
    LineTooLongError(std::size_t line, std::size_t limit, std::string_view format)
        : std::runtime_error(std::string(format) + " line " + std::to_string(line) + " is longer than " +
                             std::to_string(limit) + " bytes"),
          line_(line), limit_(limit)
    {
    }
    std::size_t line() const noexcept { return line_; }
    std::size_t limit() const noexcept { return limit_; }

private:
    std::size_t line_, limit_;
};

namespace detail {

// Lines of text with a bounded length, numbered from 1 (the rules of reference
// io/line_reader.cpp): a line ends at '\n' or at the end of the input, one '\r'
// before the end is dropped, and what remains may hold at most `limit` bytes, so
// no more than limit + 1 bytes are ever held. `format` names the input in
// messages ("OBJ", "scene").
class LineReader {
public:
    LineReader(std::istream& in, std::size_t limit, std::string_view format)
        : in_(in), limit_(limit), format_(format)
    {
        if (in.fail()) throw std::runtime_error(format_ + " input stream is not readable");
        if (limit > std::size_t(std::numeric_limits<std::streamsize>::max()) - 2)
            throw std::invalid_argument(format_ + " line length limit is too large");
    }
    // The next line into `line`; false at the end of the input.
    bool next(std::string& line)
    {
        line.clear();
        if (done_) return false;
        bool any = false, newline = false;
        for (int c; (c = in_.get()) != std::char_traits<char>::eof();) {
            any = true;
            if (c == '\n') {
                newline = true;
                break;
            }
            if (line.size() > limit_) tooLong();
            line.push_back(char(c));
        }
        if (in_.bad()) {
            done_ = true;
            throw std::runtime_error("failed to read " + format_ + " input");
        }
        if (!newline) done_ = true;
        if (!any) return false;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() > limit_) tooLong();
        ++number_;
        return true;
    }
    std::size_t number() const { return number_; }

private:
    [[noreturn]] void tooLong()
    {
        done_ = true;
        throw LineTooLongError(number_ + 1, limit_, format_);
    }

    std::istream& in_;
    std::size_t limit_;
    std::string format_;
    std::size_t number_{0};
    bool done_{false};
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
    detail::LineReader lines(input, limits.max_line_length, "OBJ");
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
    } catch (const LineTooLongError& e) {
        throw ObjParseError(e.line(), "line longer than " + std::to_string(e.limit()) + " bytes");
    }
    return TriangleMesh(std::move(pos), std::move(faces), std::move(uv), std::move(uv_faces));
}

// reference geometry/obj.cpp:251-263
inline TriangleMesh readObj(const std::filesystem::path& path, const ObjLimits& limits = {})
{
    std::ifstream file(path, std::ios::binary);
    if (!file) throw std::runtime_error("cannot open OBJ file '" + path.string() + "'");
    return readObj(file, limits);
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

// The scene format, as tables. A line is a statement: a keyword, then, by the
// statement's row of kStatements, a leading argument and `key value...` fields
// in any order (each at most once), described by rows of Field. Values are typed
// by Arg; the texts that describe them and the messages a reader gives are the
// reference's (scene/scene_file.cpp), kept here as data because callers match
// on them.
namespace text {
inline std::string expected(const std::string& what) { return "expected " + what; }
inline std::string notFinite(const std::string& what, std::string_view tok)
{
    return what + " must be a finite number, got " + quoted(tok);
}
inline std::string outOfRange(const std::string& what, Real lo, Real hi, Real v)
{
    return what + " must be between " + num(lo) + " and " + num(hi) + ", got " + num(v);
}
inline std::string notPositive(const std::string& what, Real v) { return what + " must be positive, got " + num(v); }
constexpr const char* kAxes[3] = {"the x coordinate of ", "the y coordinate of ", "the z coordinate of "};
constexpr const char* kChannels[3] = {"the red channel of ", "the green channel of ", "the blue channel of "};
inline std::string badName(std::string_view tok) { return quoted(tok) + " is not a valid name"; }
inline std::string twice(std::string_view key) { return quoted(key) + " is given twice"; }
inline std::string unknownField(std::string_view key) { return "unknown field " + quoted(key); }
inline std::string missing(std::string_view key) { return "missing '" + std::string(key) + "'"; }
inline std::string unexpected(std::string_view tok) { return "unexpected " + quoted(tok); }
inline std::string unknownStatement(std::string_view kw) { return "unknown statement " + quoted(kw); }
inline std::string tooMany(std::size_t n, const char* things) { return "more than " + std::to_string(n) + " " + things; }
inline std::string undefined(const char* kind, std::string_view name)
{
    return std::string(kind) + " " + quoted(name) + " is not defined";
}
inline std::string redefined(const char* kind, std::string_view name)
{
    return std::string(kind) + " " + quoted(name) + " is already defined";
}
inline std::string nestedMix(std::string_view name)
{
    return "material " + quoted(name) + " is a mix, and a mix cannot hold another mix";
}
inline std::string unknownMaterialType(std::string_view t) { return "unknown material type " + quoted(t); }
inline std::string unknownOperation(std::string_view op) { return "unknown instance operation " + quoted(op); }
inline std::string meshPath(std::string_view p, const char* problem) { return "mesh path " + quoted(p) + problem; }
constexpr const char* kNoCamera = "the scene has no camera";
constexpr const char* kSecondCamera = "the scene already has a camera";
constexpr const char* kSecondSky = "the scene already has a sky";
constexpr const char* kFov = "vertical field of view must be in (0, 180) degrees, got ";
constexpr const char* kCameraDistance =
    "distance from camera position to target must be non-zero, finite and within roughly [1e-154, 1e154]";
constexpr const char* kCameraUp =
    "length of the camera up direction must be non-zero, finite and within roughly [1e-154, 1e154]";
constexpr const char* kCameraParallel = "camera up direction must not be parallel to the view";
constexpr const char* kConductorChannel = "a conductor channel cannot have eta 1 and k 0";
constexpr const char* kMeshNeedsFile = "meshes can only be read from a scene file, which they sit beside";
constexpr const char* kMaterialTwice = "'material' is given twice";
constexpr const char* kOutOfRangeBeforeFit = "the instance is out of range before 'fit'";
constexpr const char* kCannotFit = "the instance is too large or too small to fit";
constexpr const char* kOutOfRangeBeforeGround = "the instance is out of range before 'ground'";
constexpr const char* kInstanceOutOfRange = "mesh instance transform takes the mesh out of range";
} // namespace text

// Value ranges of the reference's reader: reflectances in [0, 1], emission up to
// 1e6, roughness in [0, 1], dielectric index in [1e-2, 1e2], conductor eta in
// [1e-3, 1e3] and k in [0, 1e3], Perlin veins in [0, 100].
constexpr Real kReflectMax = 1, kEmitMax = 1e6, kRoughMax = 1, kIorMin = 1e-2, kIorMax = 1e2, kEtaMin = 1e-3,
               kEtaMax = 1e3, kKMax = 1e3, kVeinsMax = 1e2;

// How a value is written: three numbers (a point, a vector, a colour whose
// channels lie in [lo, hi]), one number (any, positive, or in [lo, hi]), or the
// name of a material defined earlier (Basic: one that is not a mix).
enum class Arg : uint8_t { Point, Vector, Color, Number, Positive, Ranged, Material, Basic };

struct Field {
    std::string_view key;
    Arg arg;
    const char* what; // how messages name the value
    Real lo{0}, hi{0};
};

// A field's value once read.
struct Slot {
    bool given{false};
    Real v[3]{0, 0, 0};
    const Material* material{nullptr};
    Point3r point() const { return {v[0], v[1], v[2]}; }
    Vector3r vector() const { return {v[0], v[1], v[2]}; }
    Rgb color() const { return {v[0], v[1], v[2]}; }
};

// The tokens of one statement and the line it stands on.
class Cursor {
public:
    Cursor(std::string_view text, std::size_t line) : tokens_(text), line_(line) {}
    [[noreturn]] void fail(const std::string& problem) const { throw SceneParseError(line_, problem); }
    std::optional<std::string_view> next() { return tokens_.next(); }
    std::optional<std::string_view> peek() const
    {
        Tokens copy = tokens_;
        return copy.next();
    }
    std::string_view take(const std::string& what)
    {
        const auto t = next();
        if (!t) fail(text::expected(what));
        return *t;
    }
    Real real(const std::string& what)
    {
        const std::string_view t = take(what);
        const auto v = parseReal(t);
        if (!v) fail(text::notFinite(what, t));
        return *v;
    }
    std::string_view identifier(const std::string& what)
    {
        const std::string_view t = take(what);
        if (!isIdentifier(t)) fail(text::badName(t));
        return t;
    }
    void end()
    {
        if (const auto t = next()) fail(text::unexpected(*t));
    }

private:
    Tokens tokens_;
    std::size_t line_;
};

using MaterialTable = std::map<std::string, Material, std::less<>>;

// Reads one value of kind `f.arg` into `s`; `materials` resolves names.
inline void readValue(Cursor& c, const Field& f, Slot& s, const MaterialTable& materials)
{
    const std::string what = f.what;
    const auto bounded = [&](const std::string& w) {
        const Real v = c.real(w);
        if (v < f.lo || v > f.hi) c.fail(text::outOfRange(w, f.lo, f.hi, v));
        return v;
    };
    switch (f.arg) {
    case Arg::Point:
    case Arg::Vector:
        for (int k = 0; k < 3; ++k) s.v[k] = c.real(text::kAxes[k] + what);
        break;
    case Arg::Color:
        for (int k = 0; k < 3; ++k) s.v[k] = bounded(text::kChannels[k] + what);
        break;
    case Arg::Number:
        s.v[0] = c.real(what);
        break;
    case Arg::Positive:
        s.v[0] = c.real(what);
        if (!(s.v[0] > 0)) c.fail(text::notPositive(what, s.v[0]));
        break;
    case Arg::Ranged:
        s.v[0] = bounded(what);
        break;
    case Arg::Material:
    case Arg::Basic: {
        const std::string_view name = c.identifier("a material name");
        const auto it = materials.find(name);
        if (it == materials.end()) c.fail(text::undefined("material", name));
        if (f.arg == Arg::Basic && std::holds_alternative<MixMaterial>(it->second)) c.fail(text::nestedMix(name));
        s.material = &it->second;
        break;
    }
    }
    s.given = true;
}

// The fields of a statement, read to its end: each key at most once, in any order.
template <std::size_t N>
std::array<Slot, N> readFields(Cursor& c, const std::array<Field, N>& fields, const MaterialTable& materials)
{
    std::array<Slot, N> slots{};
    while (const auto key = c.next()) {
        std::size_t i = 0;
        while (i < N && fields[i].key != *key) ++i;
        if (i == N) c.fail(text::unknownField(*key));
        if (slots[i].given) c.fail(text::twice(*key));
        readValue(c, fields[i], slots[i], materials);
    }
    return slots;
}

// Slot `i` of a statement whose field table is `fields`; fails when it was not given.
template <std::size_t N>
const Slot& need(const Cursor& c, const std::array<Field, N>& fields, const std::array<Slot, N>& slots, std::size_t i)
{
    if (!slots[i].given) c.fail(text::missing(fields[i].key));
    return slots[i];
}

inline Real orDefault(const Slot& s, Real value) { return s.given ? s.v[0] : value; }

// ------------------------------------------------------------ field tables

inline const std::array<Field, 4> kCameraFields{{{"from", Arg::Point, "the camera position"},
                                                 {"at", Arg::Point, "the point the camera looks at"},
                                                 {"up", Arg::Vector, "the camera's up direction"},
                                                 {"fov", Arg::Number, "the field of view"}}};
inline const std::array<Field, 2> kSkyFields{{{"nadir", Arg::Color, "the radiance from below", 0, kEmitMax},
                                              {"zenith", Arg::Color, "the radiance from above", 0, kEmitMax}}};
inline const Field kSkyUniform{"", Arg::Color, "the sky's radiance", 0, kEmitMax};
inline const std::array<Field, 2> kSphereFields{{{"radius", Arg::Number, "a radius"},
                                                 {"material", Arg::Material, "a material"}}};
inline const std::array<Field, 1> kPointLightFields{{{"intensity", Arg::Color, "the intensity", 0, kEmitMax}}};
inline const std::array<Field, 3> kAreaLightFields{{{"radius", Arg::Number, "a radius"},
                                                    {"radiance", Arg::Color, "the radiance", 0, kEmitMax},
                                                    {"material", Arg::Material, "a material"}}};
inline const std::array<Field, 1> kDiffuseFields{{{"color", Arg::Color, "the color", 0, kReflectMax}}};
inline const std::array<Field, 3> kCheckerFields{{{"even", Arg::Color, "the even color", 0, kReflectMax},
                                                  {"odd", Arg::Color, "the odd color", 0, kReflectMax},
                                                  {"frequency", Arg::Positive, "the checker frequency"}}};
inline const std::array<Field, 4> kPerlinFields{{{"low", Arg::Color, "the low color", 0, kReflectMax},
                                                 {"high", Arg::Color, "the high color", 0, kReflectMax},
                                                 {"scale", Arg::Positive, "the noise scale"},
                                                 {"veins", Arg::Ranged, "the veins", 0, kVeinsMax}}};
inline const std::array<Field, 3> kConductorFields{{{"eta", Arg::Color, "the refractive index", kEtaMin, kEtaMax},
                                                    {"k", Arg::Color, "the absorption", 0, kKMax},
                                                    {"roughness", Arg::Ranged, "the roughness", 0, kRoughMax}}};
inline const std::array<Field, 2> kDielectricFields{
    {{"ior", Arg::Ranged, "the index of refraction", kIorMin, kIorMax},
     {"roughness", Arg::Ranged, "the roughness", 0, kRoughMax}}};
inline const std::array<Field, 5> kMixFields{{{"first", Arg::Basic, "a material"},
                                              {"second", Arg::Basic, "a material"},
                                              {"scale", Arg::Positive, "the noise scale"},
                                              {"veins", Arg::Ranged, "the veins", 0, kVeinsMax},
                                              {"threshold", Arg::Ranged, "the threshold", 0, 1}}};

inline BasicMaterial basicOf(const Material& m)
{
    if (const auto* d = std::get_if<DiffuseMaterial>(&m)) return *d;
    if (const auto* k = std::get_if<ConductorMaterial>(&m)) return *k;
    return std::get<DielectricMaterial>(m);
}

// Material types: name -> reads the type's fields and makes the material.
struct MaterialType {
    std::string_view name;
    Material (*make)(Cursor&, const MaterialTable&);
};

inline const std::array<MaterialType, 6> kMaterialTypes{{
    {"diffuse",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kDiffuseFields, t);
         return DiffuseMaterial{ConstantTexture{need(c, kDiffuseFields, s, 0).color()}};
     }},
    {"checker",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kCheckerFields, t);
         const Rgb even = need(c, kCheckerFields, s, 0).color(), odd = need(c, kCheckerFields, s, 1).color();
         return DiffuseMaterial{CheckerTexture{even, odd, need(c, kCheckerFields, s, 2).v[0]}};
     }},
    {"perlin",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kPerlinFields, t);
         const Rgb low = need(c, kPerlinFields, s, 0).color(), high = need(c, kPerlinFields, s, 1).color();
         return DiffuseMaterial{PerlinTexture{low, high, need(c, kPerlinFields, s, 2).v[0], orDefault(s[3], 0)}};
     }},
    {"conductor",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kConductorFields, t);
         const Slot& eta = need(c, kConductorFields, s, 0);
         const Slot& k = need(c, kConductorFields, s, 1);
         for (int ch = 0; ch < 3; ++ch)
             if (eta.v[ch] == 1 && k.v[ch] == 0) c.fail(text::kConductorChannel);
         return ConductorMaterial{eta.color(), k.color(), orDefault(s[2], 0)};
     }},
    {"dielectric",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kDielectricFields, t);
         return DielectricMaterial{need(c, kDielectricFields, s, 0).v[0], orDefault(s[1], 0)};
     }},
    {"mix",
     [](Cursor& c, const MaterialTable& t) -> Material {
         const auto s = readFields(c, kMixFields, t);
         const BasicMaterial first = basicOf(*need(c, kMixFields, s, 0).material);
         const BasicMaterial second = basicOf(*need(c, kMixFields, s, 1).material);
         const Real scale = need(c, kMixFields, s, 2).v[0];
         return MixMaterial{first, second, PerlinPattern{scale, orDefault(s[3], 0)}, orDefault(s[4], 0.5)};
     }},
}};

// reference camera/perspective_camera.cpp:24-110: the checks a camera must pass.
inline void checkCamera(const Cursor& c, const CameraSettings& s)
{
    const Real fov = s.vertical_fov_degrees;
    if (std::isnan(fov) || fov <= 0 || fov >= 180) c.fail(text::kFov + num(fov));
    const Real b[3] = {s.look_from.x - s.look_at.x, s.look_from.y - s.look_at.y, s.look_from.z - s.look_at.z};
    const Real up[3] = {s.up.x, s.up.y, s.up.z};
    const Real bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2], uu = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
    if (!std::isnormal(bb)) c.fail(text::kCameraDistance);
    if (!std::isnormal(uu)) c.fail(text::kCameraUp);
    Real w[3], u[3];
    for (int k = 0; k < 3; ++k) {
        w[k] = b[k] / std::sqrt(bb);
        u[k] = up[k] / std::sqrt(uu);
    }
    Real r2 = 0;
    for (int k = 0; k < 3; ++k) {
        const Real x = u[(k + 1) % 3] * w[(k + 2) % 3] - u[(k + 2) % 3] * w[(k + 1) % 3];
        r2 += x * x;
    }
    if (r2 < 1e-12) c.fail(text::kCameraParallel);
}

// The checks of reference scene_file.cpp's mesh path rules: relative, no '..',
// no control bytes or Windows device names, inside the scene's directory after
// links are resolved, an existing regular file. Returns the resolved file.
inline std::filesystem::path sandboxedMeshFile(const Cursor& c, const std::filesystem::path& base,
                                               std::string_view path)
{
    namespace fs = std::filesystem;
    if (std::any_of(path.begin(), path.end(), [](char ch) {
            const auto b = static_cast<unsigned char>(ch);
            return b < 0x20 || b == 0x7F;
        }))
        c.fail(text::meshPath(path, " must be UTF-8 without control characters"));
    const fs::path rel{std::string(path)};
    const bool climbs = std::any_of(rel.begin(), rel.end(), [](const fs::path& part) { return part == ".."; });
    if (climbs || rel.is_absolute() || rel.has_root_name() || rel.has_root_directory())
        c.fail(text::meshPath(path, " must be relative to the scene and not use '..'"));
    static constexpr std::string_view kDevices[] = {"CON", "PRN", "AUX", "NUL"};
    for (const fs::path& part : rel) {
        std::string stem = part.string();
        stem.resize(std::min(stem.find_first_of(".:"), stem.size()));
        while (!stem.empty() && stem.back() == ' ') stem.pop_back();
        std::transform(stem.begin(), stem.end(), stem.begin(), [](char ch) {
            return ch >= 'a' && ch <= 'z' ? char(ch - 32) : ch;
        });
        const bool named = std::find(std::begin(kDevices), std::end(kDevices), stem) != std::end(kDevices);
        const bool port = stem.size() == 4 && (stem.compare(0, 3, "COM") == 0 || stem.compare(0, 3, "LPT") == 0) &&
                          stem[3] >= '1' && stem[3] <= '9';
        if (named || port) c.fail(text::meshPath(path, " names a device"));
    }
    std::error_code ec;
    const fs::path root = fs::canonical(base, ec);
    if (ec) c.fail("the scene's directory cannot be resolved");
    const fs::path file = fs::weakly_canonical(root / rel, ec);
    if (ec) c.fail("cannot resolve mesh path " + quoted(path));
    const auto [r, f] = std::mismatch(root.begin(), root.end(), file.begin(), file.end());
    (void)f;
    if (r != root.end()) c.fail(text::meshPath(path, " leads outside the scene's directory"));
    const fs::file_type type = fs::status(file, ec).type();
    if (type == fs::file_type::not_found) c.fail("mesh file " + quoted(path) + " does not exist");
    if (type != fs::file_type::regular) c.fail(text::meshPath(path, " is not a regular file"));
    return file;
}

// State of one read: what has been defined so far and the budgets left.
struct SceneState {
    SceneLimits limits;
    std::optional<std::filesystem::path> base;
    ObjLimits mesh_budget;
    std::size_t ground_left;
    std::size_t objects{0};
    std::map<std::string, std::shared_ptr<const MeshGeometry>, std::less<>> meshes;
    MaterialTable materials;
    Scene scene;
    std::optional<CameraSettings> camera;
    bool sky{false};

    SceneState(const SceneLimits& l, std::optional<std::filesystem::path> b)
        : limits(l), base(std::move(b)), mesh_budget(l.mesh_limits), ground_left(l.max_ground_points)
    {
    }
    void countObject(const Cursor& c)
    {
        if (objects >= limits.max_objects) c.fail(text::tooMany(limits.max_objects, "objects"));
        ++objects;
    }
};

inline Sphere validSphere(const Cursor& c, Point3r center, Real radius)
{
    try {
        return Sphere{center, radius};
    } catch (const std::invalid_argument& e) {
        c.fail(e.what());
    }
}

inline bool finiteBox(const Point3r& lo, const Point3r& hi)
{
    for (const Real v : {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z})
        if (!std::isfinite(v)) return false;
    return true;
}

// Words that end a one-factor `scale`: the instance statement's keywords.
inline constexpr std::array<std::string_view, 6> kInstanceWords{"material", "translate", "scale",
                                                                "rotate",   "fit",       "ground"};

// Instance operations: each applies after the placement so far
// (reference Transform composition, core/transform.hpp).
struct Operation {
    std::string_view name;
    Transform (*apply)(Cursor&, SceneState&, const MeshGeometry&, const Transform&);
};

inline const std::array<Operation, 5> kOperations{{
    {"translate",
     [](Cursor& c, SceneState&, const MeshGeometry&, const Transform& p) {
         Slot s;
         readValue(c, {"", Arg::Vector, "the translation"}, s, {});
         return Transform::translation(s.vector()) * p;
     }},
    {"scale",
     [](Cursor& c, SceneState&, const MeshGeometry&, const Transform& p) {
         // One factor, or three: a second number follows unless the statement
         // ends or the next word is a keyword of the instance statement.
         const Real sx = c.real("a scale factor");
         const auto nxt = c.peek();
         const bool uniform = !nxt || std::find(kInstanceWords.begin(), kInstanceWords.end(), *nxt) !=
                                          kInstanceWords.end();
         if (uniform) return Transform::scaling({sx, sx, sx}) * p;
         const Real sy = c.real("the y scale factor");
         return Transform::scaling({sx, sy, c.real("the z scale factor")}) * p;
     }},
    {"rotate",
     [](Cursor& c, SceneState&, const MeshGeometry&, const Transform& p) {
         Slot axis;
         readValue(c, {"", Arg::Vector, "the rotation axis"}, axis, {});
         return Transform::rotation(axis.vector(), c.real("the rotation angle")) * p;
     }},
    {"fit",
     [](Cursor& c, SceneState&, const MeshGeometry& g, const Transform& p) {
         // Scales the placed mesh so its largest box extent is `size`, centred
         // on the origin.
         Slot size;
         readValue(c, {"", Arg::Positive, "the fitted size"}, size, {});
         Point3r lo, hi;
         p.applyBox(g.boundsMin(), g.boundsMax(), lo, hi);
         if (!finiteBox(lo, hi)) c.fail(text::kOutOfRangeBeforeFit);
         const Real extent = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
         const Real k = size.v[0] / extent;
         if (!std::isfinite(k) || !(k > 0)) c.fail(text::kCannotFit);
         const Vector3r to_origin{-(lo.x / 2 + hi.x / 2), -(lo.y / 2 + hi.y / 2), -(lo.z / 2 + hi.z / 2)};
         return Transform::scaling({k, k, k}) * Transform::translation(to_origin) * p;
     }},
    {"ground",
     [](Cursor& c, SceneState& st, const MeshGeometry& g, const Transform& p) {
         // Lifts the placed mesh so its lowest used point rests on y = 0.
         const auto& points = g.usedPoints();
         if (points.size() > st.ground_left)
             c.fail("'ground' would examine more than " + std::to_string(st.limits.max_ground_points) +
                    " points in the scene");
         st.ground_left -= points.size();
         Real lowest = std::numeric_limits<Real>::infinity();
         for (const Point3r& q : points) {
             const Real y = p.apply(q).y;
             if (!std::isfinite(y)) c.fail(text::kOutOfRangeBeforeGround);
             lowest = std::min(lowest, y);
         }
         return Transform::translation({0, -lowest, 0}) * p;
     }},
}};

// Statements: keyword -> reader.
struct StatementKind {
    std::string_view keyword;
    void (*read)(Cursor&, SceneState&);
};

inline void readCameraStatement(Cursor& c, SceneState& st)
{
    if (st.camera) c.fail(text::kSecondCamera);
    const auto s = readFields(c, kCameraFields, st.materials);
    CameraSettings cam{};
    cam.look_from = need(c, kCameraFields, s, 0).point();
    cam.look_at = need(c, kCameraFields, s, 1).point();
    if (s[2].given) cam.up = s[2].vector();
    cam.vertical_fov_degrees = orDefault(s[3], cam.vertical_fov_degrees);
    checkCamera(c, cam);
    st.camera = cam;
}

inline void readSkyStatement(Cursor& c, SceneState& st)
{
    if (st.sky) c.fail(text::kSecondSky);
    // Either one radiance for the whole sky, or `nadir` and `zenith` fields.
    const auto first = c.peek();
    if (first && (*first == kSkyFields[0].key || *first == kSkyFields[1].key)) {
        const auto s = readFields(c, kSkyFields, st.materials);
        const Rgb nadir = need(c, kSkyFields, s, 0).color();
        st.scene.setSky(SkyLight{nadir, need(c, kSkyFields, s, 1).color()});
    } else {
        Slot s;
        readValue(c, kSkyUniform, s, st.materials);
        c.end();
        st.scene.setSky(SkyLight::uniform(s.color()));
    }
    st.sky = true;
}

inline void readMaterialStatement(Cursor& c, SceneState& st)
{
    if (st.materials.size() >= st.limits.max_materials) c.fail(text::tooMany(st.limits.max_materials, "materials"));
    const std::string_view name = c.identifier("a material name");
    if (st.materials.count(name)) c.fail(text::redefined("material", name));
    const std::string_view type = c.take("a material type");
    const auto kind = std::find_if(kMaterialTypes.begin(), kMaterialTypes.end(),
                                   [&](const MaterialType& m) { return m.name == type; });
    if (kind == kMaterialTypes.end()) c.fail(text::unknownMaterialType(type));
    Material m = kind->make(c, st.materials);
    st.materials.emplace(std::string(name), std::move(m));
}

inline void readSphereStatement(Cursor& c, SceneState& st)
{
    st.countObject(c);
    Slot center;
    readValue(c, {"", Arg::Point, "the sphere center"}, center, st.materials);
    const auto s = readFields(c, kSphereFields, st.materials);
    const Sphere sphere = validSphere(c, center.point(), need(c, kSphereFields, s, 0).v[0]);
    st.scene.add(sphere, *need(c, kSphereFields, s, 1).material);
}

inline void readPointLightStatement(Cursor& c, SceneState& st)
{
    st.countObject(c);
    Slot at;
    readValue(c, {"", Arg::Point, "the light position"}, at, st.materials);
    const auto s = readFields(c, kPointLightFields, st.materials);
    st.scene.addLight(PointLight{at.point(), need(c, kPointLightFields, s, 0).color()});
}

inline void readAreaLightStatement(Cursor& c, SceneState& st)
{
    st.countObject(c);
    Slot center;
    readValue(c, {"", Arg::Point, "the light center"}, center, st.materials);
    const auto s = readFields(c, kAreaLightFields, st.materials);
    const Sphere sphere = validSphere(c, center.point(), need(c, kAreaLightFields, s, 0).v[0]);
    const Rgb radiance = need(c, kAreaLightFields, s, 1).color();
    st.scene.addAreaLight(sphere, radiance, s[2].given ? *s[2].material : Material{DiffuseMaterial{}});
}

inline void readMeshStatement(Cursor& c, SceneState& st)
{
    if (!st.base) c.fail(text::kMeshNeedsFile);
    if (st.meshes.size() >= st.limits.max_meshes) c.fail(text::tooMany(st.limits.max_meshes, "meshes"));
    const std::string_view name = c.identifier("a mesh name");
    if (st.meshes.count(name)) c.fail(text::redefined("mesh", name));
    const std::string_view path = c.take("a mesh file");
    const bool smooth = c.peek() == std::optional<std::string_view>{"smooth"};
    if (smooth) (void)c.next();
    c.end();
    const std::filesystem::path file = sandboxedMeshFile(c, *st.base, path);
    std::ifstream in(file, std::ios::binary);
    if (!in) c.fail("cannot read mesh file " + quoted(path));
    std::optional<TriangleMesh> mesh;
    try {
        mesh.emplace(readObj(in, st.mesh_budget));
    } catch (const ObjParseError& e) {
        c.fail("mesh " + quoted(path) + ": " + e.what());
    } catch (const std::out_of_range& e) {
        c.fail("mesh " + quoted(path) + ": " + e.what());
    } catch (const std::runtime_error&) {
        c.fail("cannot read mesh file " + quoted(path));
    }
    std::shared_ptr<const MeshGeometry> geometry;
    try {
        geometry = std::make_shared<const MeshGeometry>(*mesh, smooth ? MeshShading::Smooth : MeshShading::Flat);
    } catch (const std::invalid_argument&) {
        c.fail("mesh " + quoted(path) + " has no triangles");
    }
    // Meshes of one scene share the OBJ budgets.
    st.mesh_budget.max_vertices -= mesh->positions().size();
    st.mesh_budget.max_texture_coordinates -= mesh->textureCoordinates().size();
    st.mesh_budget.max_triangles -= mesh->faces().size();
    st.meshes.emplace(std::string(name), std::move(geometry));
}

inline void readInstanceStatement(Cursor& c, SceneState& st)
{
    st.countObject(c);
    const std::string_view mesh_name = c.identifier("a mesh name");
    const auto mesh = st.meshes.find(mesh_name);
    if (mesh == st.meshes.end()) c.fail(text::undefined("mesh", mesh_name));
    // `material NAME` anywhere, and placement operations applied in order.
    Slot material;
    Transform placement;
    std::size_t ops = 0;
    while (const auto word = c.next()) {
        if (*word == "material") {
            if (material.given) c.fail(text::kMaterialTwice);
            readValue(c, {"material", Arg::Material, "a material"}, material, st.materials);
            continue;
        }
        if (++ops > st.limits.max_instance_operations)
            c.fail(text::tooMany(st.limits.max_instance_operations, "operations"));
        const auto op = std::find_if(kOperations.begin(), kOperations.end(),
                                     [&](const Operation& o) { return o.name == *word; });
        if (op == kOperations.end()) c.fail(text::unknownOperation(*word));
        try {
            placement = op->apply(c, st, *mesh->second, placement);
        } catch (const std::invalid_argument& e) {
            c.fail(e.what());
        }
    }
    if (!material.given) c.fail(text::missing("material"));
    Point3r lo, hi;
    placement.applyBox(mesh->second->boundsMin(), mesh->second->boundsMax(), lo, hi);
    if (!finiteBox(lo, hi)) c.fail(text::kInstanceOutOfRange);
    st.scene.add(MeshInstance{mesh->second, placement}, *material.material);
}

inline const std::array<StatementKind, 8> kStatements{{{"camera", readCameraStatement},
                                                       {"sky", readSkyStatement},
                                                       {"material", readMaterialStatement},
                                                       {"sphere", readSphereStatement},
                                                       {"point_light", readPointLightStatement},
                                                       {"area_light", readAreaLightStatement},
                                                       {"mesh", readMeshStatement},
                                                       {"instance", readInstanceStatement}}};

// The reader of reference scene/scene_file.cpp:443-1114, by behaviour: lines of
// at most max_line_length bytes, '#' starts a comment, one statement per line.
inline SceneDescription readSceneFrom(std::istream& input, std::optional<std::filesystem::path> base,
                                      const SceneLimits& limits)
{
    if (base && base->empty()) base = ".";
    SceneState st(limits, std::move(base));
    LineReader lines(input, limits.max_line_length, "scene");
    std::string line;
    try {
        while (lines.next(line)) {
            Cursor c(std::string_view(line).substr(0, line.find('#')), lines.number());
            const auto keyword = c.next();
            if (!keyword) continue;
            const auto kind = std::find_if(kStatements.begin(), kStatements.end(),
                                           [&](const StatementKind& k) { return k.keyword == *keyword; });
            if (kind == kStatements.end()) c.fail(text::unknownStatement(*keyword));
            kind->read(c, st);
        }
    } catch (const LineTooLongError& e) {
        throw SceneParseError(e.line(), "line longer than " + std::to_string(e.limit()) + " bytes");
    }
    if (!st.camera) throw SceneParseError(0, text::kNoCamera);
    return {std::move(st.scene), *st.camera};
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
