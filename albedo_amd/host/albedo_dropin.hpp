// C++ drop-in for the reference renderer's scene and render API, backed by the
// MI355X library through its C ABI (include/albedo_gpu.h).
//
// Mirrors, by name and meaning:
//   Transform                      core/transform.hpp:38-93
//   Sphere, Triangle, TriangleMesh geometry/sphere.hpp, geometry/triangle.hpp,
//                                  geometry/triangle_mesh.hpp
//   MeshGeometry, MeshShading      geometry/mesh_geometry.hpp:21-66
//   MeshInstance                   geometry/mesh_instance.hpp:18-66
//   Constant/Checker/PerlinTexture texture/*.hpp
//   Diffuse/Conductor/Dielectric/MixMaterial, PerlinPattern  material/*.hpp
//   PointLight, SkyLight           light/point_light.hpp, light/sky_light.hpp
//   Scene                          scene/scene.hpp:31-98 (add of a shape or a TriangleMesh,
//                                  addAreaLight, addLight, setSky, sky, shapeCount,
//                                  intersect, occluded)
//   Ray, Interval, HitRecord       core/ray.hpp, core/interval.hpp, scene/hit_record.hpp
//   CameraSettings, PerspectiveCamera  camera/perspective_camera.hpp:16-53
//   PathEstimator                  estimator/path_estimator.hpp:22-38
//   RenderJob, renderImage         render/render_image.hpp:17-49
//   PixelBuffer                    image/pixel_buffer.hpp
// Error behaviour follows the reference: std::invalid_argument for invalid
// arguments (AB_INVALID_ARGUMENT), std::bad_alloc for AB_OUT_OF_MEMORY, and
// std::runtime_error for device failures.
#pragma once

#include "albedo_gpu.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace albedo {

using Real = double;

struct Vector3r {
    Real x{}, y{}, z{};
};
struct Point3r {
    Real x{}, y{}, z{};
};
struct Point2r {
    Real x{}, y{};
};
struct Point2i {
    int x{}, y{};
};
struct Normal3r {
    Real x{}, y{}, z{};
};

// A ray and a parameter range (reference core/ray.hpp:13-19, core/interval.hpp:15-40).
struct Ray {
    Point3r origin{};
    Vector3r direction{};
    Point3r at(Real t) const noexcept
    {
        return {origin.x + direction.x * t, origin.y + direction.y * t, origin.z + direction.z * t};
    }
};
struct Interval {
    Real min{std::numeric_limits<Real>::infinity()};
    Real max{-std::numeric_limits<Real>::infinity()};
    bool isEmpty() const noexcept { return min > max; }
    bool contains(Real x) const noexcept { return min <= x && x <= max; }
    bool surrounds(Real x) const noexcept { return min < x && x < max; }
};

// The nearest surface a ray hits (reference scene/hit_record.hpp:13-33): normals
// face against the ray, `object` counts objects as the reference adds them.
struct HitRecord {
    Real t{};
    Point3r point{};
    Normal3r normal{};
    Normal3r shading_normal{};
    bool front_face{};
    Point2r uv{};
    std::size_t object{};
};
struct Rgb {
    Real r{}, g{}, b{};
};

namespace detail {

inline void check(ab_status st)
{
    if (st == AB_OK) return;
    const char* m = ab_last_error(nullptr);
    const std::string msg = m ? m : "";
    if (st == AB_INVALID_ARGUMENT || st == AB_PARSE_ERROR) throw std::invalid_argument(msg);
    if (st == AB_OUT_OF_MEMORY) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

// The process-wide library context on HIP device 0, created on first use.
inline ab_context* context()
{
    static std::once_flag once;
    static ab_context* ctx = nullptr;
    std::call_once(once, [] { check(ab_context_create(0, &ctx)); });
    return ctx;
}

using Mat3 = std::array<std::array<Real, 3>, 3>;

inline Mat3 identity() { return {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }

inline Mat3 mul(const Mat3& a, const Mat3& b)
{
    Mat3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Real s = 0;
            for (int k = 0; k < 3; ++k) s += a[i][k] * b[k][j];
            r[i][j] = s;
        }
    return r;
}

inline std::array<Real, 3> apply(const Mat3& m, const std::array<Real, 3>& v)
{
    return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2], m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}

} // namespace detail

// Affine map with its exact inverse (reference core/transform.cpp:93-167).
class Transform {
public:
    Transform() = default;

    static Transform translation(Vector3r v)
    {
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
            throw std::invalid_argument("transform translation component is not finite");
        Transform t;
        t.off_ = {v.x, v.y, v.z};
        t.ioff_ = {-v.x, -v.y, -v.z};
        return t;
    }

    static Transform scaling(Vector3r f)
    {
        Transform t;
        const Real s[3] = {f.x, f.y, f.z};
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(s[k]) || s[k] == 0 || !std::isfinite(1 / s[k]))
                throw std::invalid_argument("transform scale factor must be finite and non-zero with a finite reciprocal");
            t.lin_[k][k] = s[k];
            t.ilin_[k][k] = 1 / s[k];
        }
        return t;
    }

    static Transform rotation(Vector3r axis, Real degrees)
    {
        const Real l2 = axis.x * axis.x + axis.y * axis.y + axis.z * axis.z;
        if (!std::isfinite(degrees)) throw std::invalid_argument("transform rotation angle is not finite");
        if (!std::isnormal(l2)) throw std::invalid_argument("transform rotation axis has no usable direction");
        const Real l = std::sqrt(l2), x = axis.x / l, y = axis.y / l, z = axis.z / l;
        const Real r = degrees * (3.14159265358979323846 / 180);
        const Real c = std::cos(r), s = std::sin(r), t = 1 - c;
        Transform out;
        out.lin_ = {{{t * x * x + c, t * x * y - s * z, t * x * z + s * y},
                     {t * x * y + s * z, t * y * y + c, t * y * z - s * x},
                     {t * x * z - s * y, t * y * z + s * x, t * z * z + c}}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) out.ilin_[i][j] = out.lin_[j][i];
        return out;
    }

    // `*this` applied after `first`.
    Transform operator*(const Transform& first) const
    {
        Transform r;
        r.lin_ = detail::mul(lin_, first.lin_);
        const auto o = detail::apply(lin_, first.off_);
        r.off_ = {o[0] + off_[0], o[1] + off_[1], o[2] + off_[2]};
        r.ilin_ = detail::mul(first.ilin_, ilin_);
        const auto io = detail::apply(first.ilin_, ioff_);
        r.ioff_ = {io[0] + first.ioff_[0], io[1] + first.ioff_[1], io[2] + first.ioff_[2]};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j)
                if (!std::isfinite(r.lin_[i][j]) || !std::isfinite(r.ilin_[i][j]))
                    throw std::invalid_argument("composed transform has a non-finite entry");
            if (!std::isfinite(r.off_[i]) || !std::isfinite(r.ioff_[i]))
                throw std::invalid_argument("composed transform has a non-finite entry");
        }
        return r;
    }

    Transform inverse() const
    {
        Transform r;
        r.lin_ = ilin_;
        r.off_ = ioff_;
        r.ilin_ = lin_;
        r.ioff_ = off_;
        return r;
    }

    Point3r apply(Point3r p) const
    {
        const auto v = detail::apply(lin_, {p.x, p.y, p.z});
        return {v[0] + off_[0], v[1] + off_[1], v[2] + off_[2]};
    }

    Vector3r apply(Vector3r d) const
    {
        const auto v = detail::apply(lin_, {d.x, d.y, d.z});
        return {v[0], v[1], v[2]};
    }

    // Normals map by the transpose of the inverse (reference core/transform.cpp:202-205).
    Normal3r apply(Normal3r n) const
    {
        Real r[3];
        for (int k = 0; k < 3; ++k) r[k] = ilin_[0][k] * n.x + ilin_[1][k] * n.y + ilin_[2][k] * n.z;
        return {r[0], r[1], r[2]};
    }

    Ray apply(const Ray& ray) const { return {apply(ray.origin), apply(ray.direction)}; }

    // Arvo's transformed box (reference core/transform.cpp:118-132, 212-225).
    void applyBox(const Point3r& lo, const Point3r& hi, Point3r& out_lo, Point3r& out_hi) const
    {
        const Real l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
        Real mn[3], mx[3];
        for (int r = 0; r < 3; ++r) {
            for (int k = 0; k < 3; ++k) {
                const Real a = lin_[r][k] * l[k], b = lin_[r][k] * h[k];
                mn[r] = k == 0 ? std::min(a, b) : mn[r] + std::min(a, b);
                mx[r] = k == 0 ? std::max(a, b) : mx[r] + std::max(a, b);
            }
            mn[r] += off_[r];
            mx[r] += off_[r];
        }
        out_lo = {mn[0], mn[1], mn[2]};
        out_hi = {mx[0], mx[1], mx[2]};
    }

    void rows(double* fwd, double* inv) const
    {
        for (int r = 0; r < 3; ++r) {
            for (int k = 0; k < 3; ++k) {
                fwd[4 * r + k] = lin_[r][k];
                inv[4 * r + k] = ilin_[r][k];
            }
            fwd[4 * r + 3] = off_[r];
            inv[4 * r + 3] = ioff_[r];
        }
    }

private:
    detail::Mat3 lin_{detail::identity()}, ilin_{detail::identity()};
    std::array<Real, 3> off_{}, ioff_{};
};

namespace detail {

using V3 = std::array<Real, 3>;
inline V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 cross(V3 a, V3 b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline Real dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 v3(Point3r p) { return {p.x, p.y, p.z}; }
inline V3 v3(Vector3r p) { return {p.x, p.y, p.z}; }

// One primitive's hit in FP64, as the reference's shapes report it: t, point,
// outward geometric and shading normals (unit), texture coordinates.
struct PrimHit {
    Real t{};
    V3 point{}, normal{}, shading{};
    Point2r uv{};
};

// Haines' sphere test with the stable roots (reference geometry/sphere.cpp:66-135)
// and its texture coordinates (:137-147).
inline std::optional<PrimHit> sphereHit(Point3r c, Real r, const Ray& ray, Interval range)
{
    const V3 d = v3(ray.direction), l = sub(v3(ray.origin), v3(c));
    const Real a = dot(d, d);
    if (!std::isnormal(a)) return std::nullopt;
    const Real h = dot(d, l), cc = dot(l, l) - r * r;
    const V3 off = {l[0] - d[0] * (h / a), l[1] - d[1] * (h / a), l[2] - d[2] * (h / a)};
    const Real disc = a * (r * r - dot(off, off));
    if (disc < 0) return std::nullopt;
    const Real q = -(h + std::copysign(std::sqrt(disc), h));
    Real t0 = 0, t1 = 0;
    if (q != 0) {
        t0 = q / a;
        t1 = cc / q;
        if (t0 > t1) std::swap(t0, t1);
    }
    Real t = t0;
    if (!range.surrounds(t)) {
        t = t1;
        if (!range.surrounds(t)) return std::nullopt;
    }
    PrimHit hit;
    hit.t = t;
    const Point3r p = ray.at(t);
    hit.point = v3(p);
    const V3 out = sub(hit.point, v3(c));
    const Real l2 = dot(out, out);
    if (!(l2 > 0)) return std::nullopt;
    const Real il = 1 / std::sqrt(l2);
    hit.normal = {out[0] * il, out[1] * il, out[2] * il};
    hit.shading = hit.normal;
    const Real pi = 3.14159265358979323846;
    const Real lon = std::atan2(-hit.normal[2], hit.normal[0]) + pi;
    const Real lat = std::acos(std::clamp(-hit.normal[1], Real(-1), Real(1)));
    hit.uv = {lon / (2 * pi), lat / pi};
    return hit;
}

// Woop, Benthin and Wald's watertight test in FP64 (reference
// geometry/triangle.cpp:95-176): the ray moved to the origin and sheared onto +z;
// the point is interpolated from the corners, which keeps it on the plane.
inline std::optional<PrimHit> triangleHit(const std::array<V3, 3>& p, const std::array<Point2r, 3>& uv,
                                          const Ray& ray, Interval range)
{
    const V3 d = v3(ray.direction), o = v3(ray.origin);
    if (!std::isnormal(dot(d, d))) return std::nullopt;
    const V3 m = {std::abs(d[0]), std::abs(d[1]), std::abs(d[2])};
    const int kz = m[0] > m[1] ? (m[0] > m[2] ? 0 : 2) : (m[1] > m[2] ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < 0) std::swap(kx, ky);
    const Real sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1 / d[kz];
    V3 q[3];
    for (int i = 0; i < 3; ++i) {
        const V3 off = sub(p[i], o);
        q[i] = {off[kx] - sx * off[kz], off[ky] - sy * off[kz], sz * off[kz]};
    }
    auto edge = [](const V3& a, const V3& b) { return b[0] * a[1] - b[1] * a[0]; };
    const Real u = edge(q[1], q[2]), v = edge(q[2], q[0]), w = edge(q[0], q[1]);
    const bool inside = (u >= 0 && v >= 0 && w >= 0) || (u <= 0 && v <= 0 && w <= 0);
    if (!inside) return std::nullopt;
    const Real det = u + v + w;
    if (det == 0) return std::nullopt;
    const Real t = (u * q[0][2] + v * q[1][2] + w * q[2][2]) / det;
    if (!range.surrounds(t)) return std::nullopt;
    const Real b0 = u / det, b1 = v / det, b2 = w / det;
    PrimHit hit;
    hit.t = t;
    const V3 e1 = sub(p[1], p[0]), e2 = sub(p[2], p[0]);
    for (int k = 0; k < 3; ++k) hit.point[k] = p[0][k] + e1[k] * b1 + e2[k] * b2;
    const V3 n = cross(e1, e2);
    const Real il = 1 / std::sqrt(dot(n, n));
    hit.normal = {n[0] * il, n[1] * il, n[2] * il};
    hit.shading = hit.normal;
    hit.uv = {b0 * uv[0].x + b1 * uv[1].x + b2 * uv[2].x, b0 * uv[0].y + b1 * uv[1].y + b2 * uv[2].y};
    return hit;
}

inline const std::array<Point2r, 3>& defaultUv()
{
    static const std::array<Point2r, 3> uv = {Point2r{0, 0}, Point2r{1, 0}, Point2r{1, 1}};
    return uv;
}

} // namespace detail

class Sphere {
public:
    Sphere(Point3r center, Real radius) : c_(center), r_(radius)
    {
        if (!(radius > 0) || !std::isfinite(radius) || !std::isnormal(radius * radius))
            throw std::invalid_argument("sphere radius must be positive and its square a normal number");
        if (!std::isfinite(center.x) || !std::isfinite(center.y) || !std::isfinite(center.z))
            throw std::invalid_argument("sphere center must be finite");
    }
    Point3r center() const { return c_; }
    Real radius() const { return r_; }

private:
    Point3r c_;
    Real r_;
};

// True when Triangle{p0, p1, p2} would not throw: finite coordinates and an edge
// cross product whose squared length is a normal number (reference
// geometry/triangle.cpp:198-205).
inline bool formsTriangle(Point3r p0, Point3r p1, Point3r p2) noexcept
{
    for (const Point3r& p : {p0, p1, p2})
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z)) return false;
    const Real e1[3] = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, e2[3] = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
    const Real n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    return std::isnormal(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}

// A world-space triangle with per-corner texture coordinates (reference
// geometry/triangle.hpp:22-78). The device traces it as a one-face flat mesh under
// the identity map, which intersects it exactly as the same face of any mesh.
class Triangle {
public:
    Triangle(Point3r p0, Point3r p1, Point3r p2,
             const std::array<Point2r, 3>& uv = {Point2r{0, 0}, Point2r{1, 0}, Point2r{1, 1}})
        : corners_{p0, p1, p2}, uv_(uv)
    {
        for (const Point3r& p : corners_)
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z))
                throw std::invalid_argument("triangle vertices must have finite coordinates");
        if (!formsTriangle(p0, p1, p2)) throw std::invalid_argument("triangle has no computable area");
        for (const Point2r& t : uv_)
            if (!std::isfinite(t.x) || !std::isfinite(t.y))
                throw std::invalid_argument("triangle texture coordinates must be finite");
        const Real e1[3] = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, e2[3] = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
        const Real n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
        const Real l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        normal_ = {n[0] / l, n[1] / l, n[2] / l};
    }
    const std::array<Point3r, 3>& vertices() const noexcept { return corners_; }
    const std::array<Point2r, 3>& textureCoordinates() const noexcept { return uv_; }
    Normal3r normal() const noexcept { return normal_; }

private:
    std::array<Point3r, 3> corners_;
    std::array<Point2r, 3> uv_;
    Normal3r normal_{};
};

// Positions and triangle faces, with optional per-corner texture coordinates.
class TriangleMesh {
public:
    using Face = std::array<std::size_t, 3>;
    TriangleMesh(std::vector<Point3r> positions, std::vector<Face> faces, std::vector<Point2r> uv = {},
                 std::vector<Face> uv_faces = {})
        : positions_(std::move(positions)), faces_(std::move(faces)), uv_(std::move(uv)), uv_faces_(std::move(uv_faces))
    {
        // Checks in the reference's order (triangle_mesh.cpp:60-74).
        for (const Point3r& p : positions_)
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z))
                throw std::invalid_argument("mesh position has a non-finite coordinate");
        for (const Face& f : faces_)
            for (std::size_t i : f)
                if (i >= positions_.size()) throw std::out_of_range("mesh face uses a position the mesh does not have");
        for (const Point2r& t : uv_)
            if (!std::isfinite(t.x) || !std::isfinite(t.y))
                throw std::invalid_argument("mesh texture coordinate is not finite");
        if (!uv_faces_.empty() && uv_faces_.size() != faces_.size())
            throw std::invalid_argument("mesh has a different number of texture faces and faces");
        for (const Face& f : uv_faces_)
            for (std::size_t i : f)
                if (i >= uv_.size()) throw std::out_of_range("mesh texture face uses a missing coordinate");
    }
    const std::vector<Point3r>& positions() const { return positions_; }
    const std::vector<Face>& faces() const { return faces_; }
    const std::vector<Point2r>& textureCoordinates() const { return uv_; }
    const std::vector<Face>& textureFaces() const { return uv_faces_; }

private:
    std::vector<Point3r> positions_;
    std::vector<Face> faces_;
    std::vector<Point2r> uv_;
    std::vector<Face> uv_faces_;
};

enum class MeshShading : std::uint8_t { Flat, Smooth };

class MeshGeometry {
public:
    explicit MeshGeometry(const TriangleMesh& mesh, MeshShading shading = MeshShading::Flat)
        : shading_(shading)
    {
        // Box and used points of the triangles with computable area
        // (reference triangle_mesh.cpp:76-90, mesh_geometry.cpp:30-45).
        const auto& P = mesh.positions();
        std::vector<bool> used(P.size(), false);
        for (const auto& f : mesh.faces()) {
            const Point3r a = P[f[0]], b = P[f[1]], c = P[f[2]];
            const Real e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
            const Real n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                               e1[0] * e2[1] - e1[1] * e2[0]};
            if (!std::isnormal(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])) continue;
            for (std::size_t i : f) used[i] = true;
        }
        bool any = false;
        for (std::size_t i = 0; i < P.size(); ++i) {
            if (!used[i]) continue;
            used_.push_back(P[i]);
            if (!any) {
                lo_ = hi_ = P[i];
                any = true;
            }
            lo_ = {std::min(lo_.x, P[i].x), std::min(lo_.y, P[i].y), std::min(lo_.z, P[i].z)};
            hi_ = {std::max(hi_.x, P[i].x), std::max(hi_.y, P[i].y), std::max(hi_.z, P[i].z)};
        }
        if (!any) throw std::invalid_argument("mesh geometry needs at least one triangle with area");
        pos_.reserve(mesh.positions().size() * 3);
        for (const Point3r& p : mesh.positions()) {
            pos_.push_back(float(p.x));
            pos_.push_back(float(p.y));
            pos_.push_back(float(p.z));
        }
        for (const auto& f : mesh.faces())
            for (std::size_t i : f) idx_.push_back(uint32_t(i));
        if (!mesh.textureFaces().empty()) {
            for (const auto& f : mesh.textureFaces())
                for (std::size_t i : f) {
                    uv_.push_back(float(mesh.textureCoordinates()[i].x));
                    uv_.push_back(float(mesh.textureCoordinates()[i].y));
                }
        }
        // Hit records (Scene::intersect): the faces in the order the device numbers
        // them — those whose corners, rounded to the FP32 the device stores, still
        // span an area, as the device's build decides — with FP64 corners, corner
        // texture coordinates and, for smooth shading, area-weighted unit corner
        // normals (reference triangle_mesh.cpp:92-125, over the FP64 positions).
        std::vector<detail::V3> sums(shading == MeshShading::Smooth ? P.size() : 0, detail::V3{0, 0, 0});
        for (std::size_t fi = 0; fi < mesh.faces().size(); ++fi) {
            const auto& f = mesh.faces()[fi];
            detail::V3 fp[3];
            for (int k = 0; k < 3; ++k) fp[k] = {double(float(P[f[k]].x)), double(float(P[f[k]].y)), double(float(P[f[k]].z))};
            const detail::V3 fn = detail::cross(detail::sub(fp[1], fp[0]), detail::sub(fp[2], fp[0]));
            if (std::isnormal(detail::dot(fn, fn))) kept_.push_back(uint32_t(fi));
            if (!sums.empty() && formsTriangle(P[f[0]], P[f[1]], P[f[2]])) {
                const detail::V3 an = detail::cross(detail::sub(detail::v3(P[f[1]]), detail::v3(P[f[0]])),
                                                    detail::sub(detail::v3(P[f[2]]), detail::v3(P[f[0]])));
                for (std::size_t i : f)
                    for (int k = 0; k < 3; ++k) sums[i][k] += an[k];
            }
        }
        for (uint32_t fi : kept_) {
            const auto& f = mesh.faces()[fi];
            Face face;
            for (int k = 0; k < 3; ++k) face.p[k] = detail::v3(P[f[k]]);
            face.uv = detail::defaultUv();
            if (!mesh.textureFaces().empty())
                for (int k = 0; k < 3; ++k) face.uv[k] = mesh.textureCoordinates()[mesh.textureFaces()[fi][k]];
            if (!sums.empty()) {
                const detail::V3 fn = detail::cross(detail::sub(face.p[1], face.p[0]), detail::sub(face.p[2], face.p[0]));
                const Real fl2 = detail::dot(fn, fn);
                const detail::V3 face_n = std::isnormal(fl2) ? detail::V3{fn[0] / std::sqrt(fl2), fn[1] / std::sqrt(fl2),
                                                                          fn[2] / std::sqrt(fl2)}
                                                             : detail::V3{0, 0, 0};
                for (int k = 0; k < 3; ++k) {
                    const detail::V3& sm = sums[f[k]];
                    const Real l2 = detail::dot(sm, sm);
                    face.n[k] = std::isnormal(l2) ? detail::V3{sm[0] / std::sqrt(l2), sm[1] / std::sqrt(l2),
                                                               sm[2] / std::sqrt(l2)}
                                                  : face_n;
                }
            }
            faces_.push_back(face);
        }
    }
    MeshShading shading() const { return shading_; }
    Point3r boundsMin() const { return lo_; }
    Point3r boundsMax() const { return hi_; }
    const std::vector<Point3r>& usedPoints() const { return used_; }
    const std::vector<float>& positions() const { return pos_; }
    const std::vector<uint32_t>& indices() const { return idx_; }
    const std::vector<float>& uv() const { return uv_; }
    // Faces with area in the device's numbering (hit records' prim ids).
    std::size_t keptCount() const { return faces_.size(); }

    // The FP64 hit on kept face `k` of a ray in mesh space (reference
    // geometry/mesh_geometry.cpp:77-98, with the smooth normal blended at the
    // point's barycentric coordinates, :53-64). `t_device` is the device's t for
    // a hit the FP64 test misses by rounding (at an edge): the point then lies at
    // that t, with the face's normal.
    detail::PrimHit hitFace(std::size_t k, const Ray& ray, Interval range, Real t_device) const
    {
        const Face& f = faces_[k];
        const std::array<detail::V3, 3> p = {f.p[0], f.p[1], f.p[2]};
        std::optional<detail::PrimHit> h = detail::triangleHit(p, f.uv, ray, range);
        if (!h) {
            detail::PrimHit g;
            g.t = t_device;
            g.point = detail::v3(ray.at(t_device));
            const detail::V3 n = detail::cross(detail::sub(p[1], p[0]), detail::sub(p[2], p[0]));
            const Real il = 1 / std::sqrt(detail::dot(n, n));
            g.normal = g.shading = {n[0] * il, n[1] * il, n[2] * il};
            const auto b = barycentric(p, g.point);
            g.uv = {b[0] * f.uv[0].x + b[1] * f.uv[1].x + b[2] * f.uv[2].x,
                    b[0] * f.uv[0].y + b[1] * f.uv[1].y + b[2] * f.uv[2].y};
            h = g;
        }
        if (shading_ == MeshShading::Smooth) {
            const auto b = barycentric(p, h->point);
            detail::V3 bl{};
            for (int c = 0; c < 3; ++c)
                for (int q = 0; q < 3; ++q) bl[q] += f.n[c][q] * b[c];
            const Real l2 = detail::dot(bl, bl);
            if (l2 >= 1e-12) h->shading = {bl[0] / std::sqrt(l2), bl[1] / std::sqrt(l2), bl[2] / std::sqrt(l2)};
        }
        return *h;
    }

private:
    struct Face {
        detail::V3 p[3];
        std::array<Point2r, 3> uv;
        detail::V3 n[3]{};
    };

    // Signed-area weights of `x` (reference geometry/triangle.cpp:183-196).
    static std::array<Real, 3> barycentric(const std::array<detail::V3, 3>& p, const detail::V3& x)
    {
        const detail::V3 e1 = detail::sub(p[1], p[0]), e2 = detail::sub(p[2], p[0]), off = detail::sub(x, p[0]);
        const detail::V3 an = detail::cross(e1, e2);
        const Real a2 = detail::dot(an, an);
        const Real b1 = detail::dot(detail::cross(off, e2), an) / a2, b2 = detail::dot(detail::cross(e1, off), an) / a2;
        return {1 - b1 - b2, b1, b2};
    }

    MeshShading shading_;
    Point3r lo_{}, hi_{};
    std::vector<Point3r> used_;
    std::vector<float> pos_;
    std::vector<uint32_t> idx_;
    std::vector<float> uv_;
    std::vector<uint32_t> kept_;
    std::vector<Face> faces_;
};

class MeshInstance {
public:
    MeshInstance(std::shared_ptr<const MeshGeometry> mesh, const Transform& object_to_world)
        : mesh_(std::move(mesh)), to_world_(object_to_world)
    {
        if (!mesh_) throw std::invalid_argument("mesh instance needs a mesh");
    }
    const MeshGeometry& mesh() const { return *mesh_; }
    const std::shared_ptr<const MeshGeometry>& meshPtr() const { return mesh_; }
    const Transform& transform() const { return to_world_; }

private:
    std::shared_ptr<const MeshGeometry> mesh_;
    Transform to_world_;
};

using Shape = std::variant<Sphere, Triangle, MeshInstance>;

struct ConstantTexture {
    Rgb color{};
};
struct CheckerTexture {
    Rgb even{};
    Rgb odd{};
    Real frequency{1};
};
struct PerlinTexture {
    Rgb low{};
    Rgb high{1, 1, 1};
    Real scale{1};
    Real veins{0};
};
using Texture = std::variant<ConstantTexture, CheckerTexture, PerlinTexture>;

struct DiffuseMaterial {
    Texture albedo;
};
struct ConductorMaterial {
    Rgb eta{1, 1, 1};
    Rgb k{};
    Real roughness{};
};
struct DielectricMaterial {
    Real ior{1.5};
    Real roughness{};
};
struct PerlinPattern {
    Real scale{1};
    Real veins{0};
};
using BasicMaterial = std::variant<DiffuseMaterial, ConductorMaterial, DielectricMaterial>;
struct MixMaterial {
    BasicMaterial first;
    BasicMaterial second;
    PerlinPattern pattern;
    Real threshold{0.5};
};
using Material = std::variant<DiffuseMaterial, ConductorMaterial, DielectricMaterial, MixMaterial>;

struct PointLight {
    Point3r position{};
    Rgb intensity{};
};
struct SkyLight {
    Rgb nadir{};
    Rgb zenith{};
    static SkyLight uniform(Rgb c) { return {c, c}; }
};

// Objects and lights; flattens into an ab_scene_desc when rendered.
class Scene {
public:
    void add(const Shape& shape, const Material& material) { objects_.push_back({shape, material, -1, 1, false}); }

    // Every triangle of `mesh` with a computable area, in face order, with one
    // material (reference scene/scene.cpp:33-60). They become one flat mesh under
    // the identity map: the same faces hit exactly as separate triangles would,
    // and the (t, object, face) tie order is the same because the faces keep their
    // order and stay together in the object order. A mesh with no such triangle
    // adds nothing.
    void add(const TriangleMesh& mesh, const Material& material)
    {
        std::size_t kept = 0;
        const auto& P = mesh.positions();
        for (const auto& f : mesh.faces()) kept += formsTriangle(P[f[0]], P[f[1]], P[f[2]]) ? 1 : 0;
        if (kept == 0) return;
        auto geometry = std::make_shared<const MeshGeometry>(mesh, MeshShading::Flat);
        objects_.push_back({MeshInstance{std::move(geometry), Transform{}}, material, -1, kept, true});
    }

    void addAreaLight(const Sphere& sphere, Rgb radiance, const Material& material = DiffuseMaterial{})
    {
        const int light = int(lights_.size());
        Light l;
        l.kind = AB_LIGHT_AREA;
        l.object = uint32_t(objects_.size());
        l.radiance = radiance;
        lights_.push_back(l);
        objects_.push_back({sphere, material, light, 1, false});
    }

    std::size_t addLight(const PointLight& light)
    {
        Light l;
        l.kind = AB_LIGHT_POINT;
        l.point = light;
        lights_.push_back(l);
        return lights_.size() - 1;
    }

    void setSky(const SkyLight& sky)
    {
        Light l;
        l.kind = AB_LIGHT_SKY;
        l.sky = sky;
        if (sky_) {
            lights_[*sky_] = l;
        } else {
            sky_ = lights_.size();
            lights_.push_back(l);
        }
    }

    // Objects as the reference counts them: one per shape, one per kept triangle
    // of an added TriangleMesh.
    std::size_t shapeCount() const
    {
        std::size_t n = 0;
        for (const Object& o : objects_) n += o.count;
        return n;
    }

    // The light arriving from beyond the scene; black until set (reference
    // scene/scene.hpp:57-58).
    SkyLight sky() const { return sky_ ? lights_[*sky_].sky : SkyLight{}; }

    // The nearest hit with t strictly inside `t_range` (reference
    // scene/scene.cpp:107-133): the device picks the object and primitive (FP32
    // traversal, the (t, object, prim) lexicographic minimum), then the record is
    // computed in FP64 on that primitive as the reference's shapes do: t, point,
    // normals facing against the ray, texture coordinates, and the object index
    // as the reference numbers objects (each face of an added TriangleMesh one).
    // One device query per call; the vector form below batches many rays.
    std::optional<HitRecord> intersect(const Ray& ray, Interval t_range) const
    {
        return intersect(std::vector<Ray>{ray}, t_range)[0];
    }

    // True when any object lies along `ray` with t strictly inside `t_range`
    // (reference scene/scene.cpp:157-165).
    bool occluded(const Ray& ray, Interval t_range) const { return occluded(std::vector<Ray>{ray}, t_range)[0] != 0; }

    // Many rays in one device query (no reference counterpart: the reference
    // intersects one ray per call).
    std::vector<std::optional<HitRecord>> intersect(const std::vector<Ray>& rays, Interval t_range) const
    {
        std::vector<std::optional<HitRecord>> out(rays.size());
        if (rays.empty() || !(t_range.min < t_range.max)) return out;
        const std::vector<ab_ray> r = deviceRays(rays, t_range);
        std::vector<ab_hit> h(r.size());
        detail::check(ab_trace_closest_host(device(), r.data(), h.data(), r.size()));
        std::vector<std::size_t> base(objects_.size());
        for (std::size_t k = 0, b = 0; k < objects_.size(); b += objects_[k].count, ++k) base[k] = b;
        for (std::size_t i = 0; i < rays.size(); ++i) {
            if (h[i].object == AB_MISS) continue;
            out[i] = record(rays[i], t_range, h[i], base);
        }
        return out;
    }

    // Occlusion of many rays in one device query: 1 where something lies along the ray.
    std::vector<uint8_t> occluded(const std::vector<Ray>& rays, Interval t_range) const
    {
        std::vector<uint8_t> out(rays.size(), 0);
        if (rays.empty() || !(t_range.min < t_range.max)) return out;
        const std::vector<ab_ray> r = deviceRays(rays, t_range);
        detail::check(ab_trace_any_host(device(), r.data(), out.data(), r.size()));
        return out;
    }

    // The resident device scene (built on first use, shared by content hash).
    ab_scene* device() const
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (!device_) device_.reset(build(), [](ab_scene* s) { ab_scene_release(s); });
        return device_.get();
    }

    Scene() = default;
    Scene(const Scene& o) : objects_(o.objects_), lights_(o.lights_), sky_(o.sky_) {}

private:
    struct Object {
        Shape shape;
        Material material;
        int light;
        std::size_t count; // objects the reference counts for it
        bool faces;        // an added TriangleMesh: each kept face is an object of its own
    };
    struct Light {
        uint32_t kind{};
        uint32_t object{};
        Rgb radiance{};
        PointLight point{};
        SkyLight sky{};
    };

    static std::vector<ab_ray> deviceRays(const std::vector<Ray>& rays, Interval t_range)
    {
        std::vector<ab_ray> r(rays.size());
        for (std::size_t i = 0; i < rays.size(); ++i) {
            const Ray& q = rays[i];
            r[i] = ab_ray{{float(q.origin.x), float(q.origin.y), float(q.origin.z)},
                          float(t_range.min),
                          {float(q.direction.x), float(q.direction.y), float(q.direction.z)},
                          float(t_range.max)};
        }
        return r;
    }

    // The FP64 record of the device's hit `h` (reference scene/scene.cpp:124-133).
    HitRecord record(const Ray& ray, Interval range, const ab_hit& h, const std::vector<std::size_t>& base) const
    {
        const Object& o = objects_[h.object];
        detail::PrimHit p;
        std::size_t index = base[h.object];
        if (const auto* s = std::get_if<Sphere>(&o.shape)) {
            std::optional<detail::PrimHit> q = detail::sphereHit(s->center(), s->radius(), ray, range);
            if (!q) {
                // Rounding put the FP32 hit just inside the range and the FP64 roots
                // just outside: the point at the device's t, the sphere's normal there.
                detail::PrimHit g;
                g.t = h.t;
                g.point = detail::v3(ray.at(h.t));
                const detail::V3 out = detail::sub(g.point, detail::v3(s->center()));
                const Real il = 1 / std::sqrt(detail::dot(out, out));
                g.normal = g.shading = {out[0] * il, out[1] * il, out[2] * il};
                const Real pi = 3.14159265358979323846;
                g.uv = {(std::atan2(-g.normal[2], g.normal[0]) + pi) / (2 * pi),
                        std::acos(std::clamp(-g.normal[1], Real(-1), Real(1))) / pi};
                q = g;
            }
            p = *q;
        } else if (const auto* t = std::get_if<Triangle>(&o.shape)) {
            const std::array<detail::V3, 3> c = {detail::v3(t->vertices()[0]), detail::v3(t->vertices()[1]),
                                                 detail::v3(t->vertices()[2])};
            std::optional<detail::PrimHit> q = detail::triangleHit(c, t->textureCoordinates(), ray, range);
            p = q ? *q : MeshGeometry(TriangleMesh{{t->vertices()[0], t->vertices()[1], t->vertices()[2]},
                                                   {{0, 1, 2}},
                                                   {t->textureCoordinates()[0], t->textureCoordinates()[1],
                                                    t->textureCoordinates()[2]},
                                                   {{0, 1, 2}}})
                             .hitFace(0, ray, range, h.t);
        } else {
            const auto& inst = std::get<MeshInstance>(o.shape);
            if (o.faces) index += h.prim;
            // Mesh space with the inverse map, t shared; normals back by the
            // transpose of the inverse (reference geometry/mesh_instance.cpp:32-60).
            const Transform to_object = inst.transform().inverse();
            const Ray local = to_object.apply(ray);
            p = inst.mesh().hitFace(h.prim, local, range, h.t);
            const Point3r wp = inst.transform().apply(Point3r{p.point[0], p.point[1], p.point[2]});
            p.point = detail::v3(wp);
            const Normal3r wn = inst.transform().apply(Normal3r{p.normal[0], p.normal[1], p.normal[2]});
            const Normal3r ws = inst.transform().apply(Normal3r{p.shading[0], p.shading[1], p.shading[2]});
            const detail::V3 n = {wn.x, wn.y, wn.z}, sn = {ws.x, ws.y, ws.z};
            const Real nl = std::sqrt(detail::dot(n, n)), sl = std::sqrt(detail::dot(sn, sn));
            const bool flat = p.normal == p.shading;
            p.normal = {n[0] / nl, n[1] / nl, n[2] / nl};
            p.shading = flat || !std::isnormal(sl * sl) ? p.normal : detail::V3{sn[0] / sl, sn[1] / sl, sn[2] / sl};
        }
        HitRecord r;
        r.t = p.t;
        r.point = {p.point[0], p.point[1], p.point[2]};
        r.front_face = detail::dot(detail::v3(ray.direction), p.normal) < 0;
        const Real sign = r.front_face ? 1 : -1;
        r.normal = {sign * p.normal[0], sign * p.normal[1], sign * p.normal[2]};
        r.shading_normal = {sign * p.shading[0], sign * p.shading[1], sign * p.shading[2]};
        r.uv = p.uv;
        r.object = index;
        return r;
    }

    static void putRgb(float* d, Rgb c)
    {
        d[0] = float(c.r);
        d[1] = float(c.g);
        d[2] = float(c.b);
    }

    static ab_texture_desc texture(const Texture& t)
    {
        ab_texture_desc d{};
        if (const auto* c = std::get_if<ConstantTexture>(&t)) {
            d.kind = AB_TEX_CONSTANT;
            putRgb(d.a, c->color);
        } else if (const auto* k = std::get_if<CheckerTexture>(&t)) {
            d.kind = AB_TEX_CHECKER;
            putRgb(d.a, k->even);
            putRgb(d.b, k->odd);
            d.frequency = float(k->frequency);
        } else {
            const auto& p = std::get<PerlinTexture>(t);
            d.kind = AB_TEX_PERLIN;
            putRgb(d.a, p.low);
            putRgb(d.b, p.high);
            d.frequency = float(p.scale);
            d.veins = float(p.veins);
        }
        return d;
    }

    static ab_material_desc basic(const BasicMaterial& m)
    {
        ab_material_desc d{};
        if (const auto* df = std::get_if<DiffuseMaterial>(&m)) {
            d.kind = AB_MAT_DIFFUSE;
            d.albedo = texture(df->albedo);
        } else if (const auto* c = std::get_if<ConductorMaterial>(&m)) {
            d.kind = AB_MAT_CONDUCTOR;
            putRgb(d.eta, c->eta);
            putRgb(d.k, c->k);
            d.roughness = float(c->roughness);
        } else {
            const auto& e = std::get<DielectricMaterial>(m);
            d.kind = AB_MAT_DIELECTRIC;
            d.ior = float(e.ior);
            d.roughness = float(e.roughness);
        }
        return d;
    }

    ab_scene* build() const
    {
        std::vector<ab_material_desc> mats;
        std::vector<ab_mesh_desc> meshes;
        std::vector<const MeshGeometry*> mesh_ptrs;
        std::vector<ab_object_desc> objs;
        // One-face meshes of the Triangle shapes, kept alive until the scene is created.
        std::vector<std::array<float, 9>> tri_pos;
        std::vector<std::array<float, 6>> tri_uv;
        std::size_t n_tri = 0;
        for (const Object& o : objects_) n_tri += std::holds_alternative<Triangle>(o.shape) ? 1 : 0;
        tri_pos.reserve(n_tri);
        tri_uv.reserve(n_tri);
        static const uint32_t kOneFace[3] = {0, 1, 2};
        double identity_fwd[12], identity_inv[12];
        Transform{}.rows(identity_fwd, identity_inv);
        for (const Object& o : objects_) {
            ab_object_desc od{};
            od.light = o.light;
            od.material = uint32_t(mats.size());
            if (const auto* mix = std::get_if<MixMaterial>(&o.material)) {
                mats.push_back(basic(mix->first));
                mats.push_back(basic(mix->second));
                ab_material_desc m{};
                m.kind = AB_MAT_MIX;
                m.first = od.material;
                m.second = od.material + 1;
                m.scale = float(mix->pattern.scale);
                m.veins = float(mix->pattern.veins);
                m.threshold = float(mix->threshold);
                od.material = uint32_t(mats.size());
                mats.push_back(m);
            } else {
                mats.push_back(std::visit(
                    [](const auto& b) -> ab_material_desc {
                        using T = std::decay_t<decltype(b)>;
                        if constexpr (std::is_same_v<T, MixMaterial>) return {};
                        else return basic(BasicMaterial{b});
                    },
                    o.material));
            }
            if (const auto* s = std::get_if<Sphere>(&o.shape)) {
                od.kind = AB_OBJ_SPHERE;
                od.center[0] = s->center().x;
                od.center[1] = s->center().y;
                od.center[2] = s->center().z;
                od.radius = s->radius();
            } else if (const auto* t = std::get_if<Triangle>(&o.shape)) {
                od.kind = AB_OBJ_INSTANCE;
                std::array<float, 9>& p = tri_pos.emplace_back();
                std::array<float, 6>& w = tri_uv.emplace_back();
                for (int k = 0; k < 3; ++k) {
                    const Point3r v = t->vertices()[k];
                    p[3 * k] = float(v.x);
                    p[3 * k + 1] = float(v.y);
                    p[3 * k + 2] = float(v.z);
                    w[2 * k] = float(t->textureCoordinates()[k].x);
                    w[2 * k + 1] = float(t->textureCoordinates()[k].y);
                }
                ab_mesh_desc md{};
                md.positions = p.data();
                md.num_vertices = 3;
                md.indices = kOneFace;
                md.num_faces = 1;
                md.uv = w.data();
                md.smooth = 0;
                od.mesh = uint32_t(meshes.size());
                meshes.push_back(md);
                mesh_ptrs.push_back(nullptr);
                std::memcpy(od.to_world, identity_fwd, sizeof identity_fwd);
                std::memcpy(od.to_object, identity_inv, sizeof identity_inv);
            } else {
                const auto& inst = std::get<MeshInstance>(o.shape);
                od.kind = AB_OBJ_INSTANCE;
                const MeshGeometry* g = &inst.mesh();
                std::size_t mi = 0;
                while (mi < mesh_ptrs.size() && mesh_ptrs[mi] != g) ++mi;
                if (mi == mesh_ptrs.size()) {
                    mesh_ptrs.push_back(g);
                    ab_mesh_desc md{};
                    md.positions = g->positions().data();
                    md.num_vertices = uint32_t(g->positions().size() / 3);
                    md.indices = g->indices().data();
                    md.num_faces = uint32_t(g->indices().size() / 3);
                    md.uv = g->uv().empty() ? nullptr : g->uv().data();
                    md.smooth = g->shading() == MeshShading::Smooth;
                    meshes.push_back(md);
                }
                od.mesh = uint32_t(mi);
                inst.transform().rows(od.to_world, od.to_object);
            }
            objs.push_back(od);
        }
        std::vector<ab_light_desc> lights;
        for (const Light& l : lights_) {
            ab_light_desc d{};
            d.kind = l.kind;
            d.object = l.object;
            d.position[0] = l.point.position.x;
            d.position[1] = l.point.position.y;
            d.position[2] = l.point.position.z;
            putRgb(d.intensity, l.point.intensity);
            putRgb(d.radiance, l.radiance);
            putRgb(d.nadir, l.sky.nadir);
            putRgb(d.zenith, l.sky.zenith);
            lights.push_back(d);
        }
        ab_scene_desc desc{meshes.data(), uint32_t(meshes.size()), objs.data(), uint32_t(objs.size()),
                           mats.data(), uint32_t(mats.size()), lights.data(), uint32_t(lights.size()), 0};
        ab_scene* s = nullptr;
        detail::check(ab_scene_create(detail::context(), &desc, &s));
        return s;
    }

    std::vector<Object> objects_;
    std::vector<Light> lights_;
    std::optional<std::size_t> sky_;
    mutable std::mutex mu_;
    mutable std::shared_ptr<ab_scene> device_;
};

struct CameraSettings {
    Point3r look_from{};
    Point3r look_at{};
    Vector3r up{0.0, 1.0, 0.0};
    Real vertical_fov_degrees{60.0};
    Real aspect_ratio{16.0 / 9.0};
};

class PerspectiveCamera {
public:
    explicit PerspectiveCamera(const CameraSettings& s) : s_(s)
    {
        if (std::isnan(s.vertical_fov_degrees) || s.vertical_fov_degrees <= 0 || s.vertical_fov_degrees >= 180)
            throw std::invalid_argument("vertical field of view must be in (0, 180) degrees");
        if (!(s.aspect_ratio > 0) || !std::isnormal(s.aspect_ratio))
            throw std::invalid_argument("aspect ratio must be a positive normal number");
    }
    const CameraSettings& settings() const { return s_; }

private:
    CameraSettings s_;
};

// The estimators of reference estimator/radiance_estimator.hpp:36-38; each is a
// mode of the device path tracer (ab_render_params::estimator).
struct NormalsEstimator {};
struct ScatterOnlyEstimator {
    int max_scatter_count{32};
};
struct PathEstimator {
    int max_depth{64};
    int roulette_start{3};
};
using RadianceEstimator = std::variant<NormalsEstimator, ScatterOnlyEstimator, PathEstimator>;

// Radiance samples accumulated per pixel: a sum and a sample count each
// (reference image/pixel_buffer.hpp:17-57, pixel_buffer.cpp:9-45). renderImage
// adds the device's per-pixel sums, which are exact fixed-point integers
// (ab_render_sums), so a pixel rendered in several calls averages exactly as
// one ab_sums_merge of those renders does.
class PixelBuffer {
public:
    static constexpr std::size_t max_pixel_count{std::size_t{1} << 28};

    PixelBuffer(int width, int height) : w_(width), h_(height)
    {
        if (!isValidSize(width, height))
            throw std::invalid_argument("invalid image size " + std::to_string(width) + "x" + std::to_string(height) +
                                        ": dimensions must be positive and the image at most " +
                                        std::to_string(max_pixel_count) + " pixels");
        pixels_.resize(std::size_t(width) * std::size_t(height));
    }
    static constexpr bool isValidSize(int width, int height) noexcept
    {
        if (width <= 0 || height <= 0) return false;
        return std::size_t(width) <= max_pixel_count / std::size_t(height);
    }
    int width() const noexcept { return w_; }
    int height() const noexcept { return h_; }
    // Precondition: `pixel` lies inside the image.
    void addSample(Point2i pixel, Rgb sample) noexcept
    {
        Accumulator& a = pixels_[index(pixel)];
        a.sum.r += sample.r;
        a.sum.g += sample.g;
        a.sum.b += sample.b;
        ++a.sample_count;
    }
    // The mean of the pixel's samples, or black if it has none.
    Rgb average(Point2i pixel) const noexcept
    {
        const Accumulator& a = pixels_[index(pixel)];
        if (a.sample_count == 0) return {};
        const Real n = Real(a.sample_count);
        return {a.sum.r / n, a.sum.g / n, a.sum.b / n};
    }
    // Adds per-pixel sums in the wire format (4 words per pixel, row-major):
    // the fixed-point sums are exact multiples of 2^-AB_SUMS_FRACTION_BITS, so
    // their FP64 values add exactly while below 2^(53-AB_SUMS_FRACTION_BITS).
    void addSums(const uint64_t* words) noexcept
    {
        const Real scale = Real(uint64_t(1) << AB_SUMS_FRACTION_BITS);
        for (std::size_t p = 0; p < pixels_.size(); ++p) {
            const uint64_t* w = words + 4 * p;
            Accumulator& a = pixels_[p];
            a.sum.r += Real(w[0]) / scale;
            a.sum.g += Real(w[1]) / scale;
            a.sum.b += Real(w[2]) / scale;
            a.sample_count += int64_t(w[3]);
        }
    }

private:
    struct Accumulator {
        Rgb sum{};
        int64_t sample_count{0};
    };
    std::size_t index(Point2i p) const noexcept { return std::size_t(p.y) * std::size_t(w_) + std::size_t(p.x); }

    int w_, h_;
    std::vector<Accumulator> pixels_;
};

// reference render/render_image.hpp:17-32, with the same defaults.
struct RenderJob {
    std::reference_wrapper<const PerspectiveCamera> camera;
    std::reference_wrapper<const Scene> scene;
    int samples_per_axis{1};
    std::uint64_t seed{0};
    RadianceEstimator estimator{ScatterOnlyEstimator{}};
    int thread_count{1}; // accepted for source compatibility; the GPU ignores it
};

// Renders `job` into `image`, adding samples_per_axis^2 samples to every pixel
// (reference render/render_image.cpp:58-77).
inline void renderImage(const RenderJob& job, PixelBuffer& image)
{
    if (job.samples_per_axis <= 0)
        throw std::invalid_argument("samples per axis must be positive, got " + std::to_string(job.samples_per_axis));
    if (job.thread_count <= 0)
        throw std::invalid_argument("thread count must be positive, got " + std::to_string(job.thread_count));
    ab_render_params p{};
    if (const auto* path = std::get_if<PathEstimator>(&job.estimator)) {
        if (path->max_depth < 0 || path->roulette_start < 0)
            throw std::invalid_argument("path limits must not be negative");
        p.estimator = AB_ESTIMATOR_PATH;
        p.max_depth = uint32_t(path->max_depth);
        p.roulette_start = uint32_t(path->roulette_start);
    } else if (const auto* scatter = std::get_if<ScatterOnlyEstimator>(&job.estimator)) {
        if (scatter->max_scatter_count < 0) throw std::invalid_argument("scatter count limit must not be negative");
        p.estimator = AB_ESTIMATOR_SCATTER_ONLY;
        p.max_depth = uint32_t(scatter->max_scatter_count);
    } else {
        p.estimator = AB_ESTIMATOR_NORMALS;
    }
    const CameraSettings& cs = job.camera.get().settings();
    ab_camera cam{};
    const Real from[3] = {cs.look_from.x, cs.look_from.y, cs.look_from.z};
    const Real at[3] = {cs.look_at.x, cs.look_at.y, cs.look_at.z};
    const Real up[3] = {cs.up.x, cs.up.y, cs.up.z};
    for (int k = 0; k < 3; ++k) {
        cam.look_from[k] = from[k];
        cam.look_at[k] = at[k];
        cam.up[k] = up[k];
    }
    cam.vfov_degrees = cs.vertical_fov_degrees;
    cam.aspect = cs.aspect_ratio;
    p.width = uint32_t(image.width());
    p.height = uint32_t(image.height());
    p.samples_per_axis = uint32_t(job.samples_per_axis);
    p.seed = job.seed;
    std::vector<uint64_t> sums(std::size_t(image.width()) * std::size_t(image.height()) * 4, 0);
    detail::check(ab_render_sums(job.scene.get().device(), &cam, &p, sums.data(), nullptr));
    image.addSums(sums.data());
}

inline int availableThreadCount() noexcept { return 1; }

} // namespace albedo
