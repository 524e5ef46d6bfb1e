// Oracle core: a CPU restatement of the reference renderer's geometry, sampling
// and acceleration code, used ONLY by tests, smoke() and bench.py's cpu_baseline
// leg as the checker. Nothing in the product (albedo_amd/) includes or links it.
//
// Templated on the real type R: R = double restates the reference (which is FP64
// throughout, src/albedo/core/real.hpp:10); R = float is the FP32 mirror whose
// operation order is the one the GPU kernels use, so GPU hit ids and t can be
// compared bit for bit. Compile with -ffp-contract=off (the reference builds
// triangle.cpp and transform.cpp that way, src/CMakeLists.txt:118-124).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

// Work counters of the ordered traversal on the calling thread: box tests and
// primitive visits. Only used to report how many bytes a traversal touches.
struct TraversalCounts {
    uint64_t nodes{0};
    uint64_t prims{0};
};
inline thread_local TraversalCounts tl_counts;

// ------------------------------------------------------------------ vectors

template <class R>
struct V3 {
    R x{}, y{}, z{};
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R lengthSquared(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return std::sqrt(lengthSquared(a)); }
template <class R> inline V3<R> normalized(V3<R> a) { return a / length(a); }
template <class R> inline bool isFiniteV(V3<R> a)
{
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
template <class T, class R> inline V3<T> cast(V3<R> a) { return {T(a.x), T(a.y), T(a.z)}; }

using V3d = V3<double>;
using V3f = V3<float>;

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 0.31830988618379067154;

// ---------------------------------------------------------------- sampling

// PCG32 (PCG-XSH-RR 64/32), O'Neill 2014; reference sampling/pcg32.hpp:30-57.
struct Pcg32 {
    static constexpr uint64_t kMult = 6364136223846793005ULL;
    uint64_t state{0};
    uint64_t inc{1};
    Pcg32(uint64_t seed, uint64_t stream) : state(0), inc((stream << 1u) | 1u)
    {
        (void)next();
        state += seed;
        (void)next();
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * kMult + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // Jump ahead by `delta` steps (O'Neill's pcg_advance_lcg_64, after Brown,
    // "Random Number Generation with Arbitrary Stride", 1994).
    void advance(uint64_t delta)
    {
        uint64_t acc_mult = 1, acc_plus = 0, cur_mult = kMult, cur_plus = inc;
        while (delta > 0) {
            if (delta & 1u) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1u;
        }
        state = acc_mult * state + acc_plus;
    }
};

// SplitMix64 finaliser; reference sampling/sample_generator.hpp:17-22.
inline uint64_t mixBits(uint64_t v)
{
    v = (v ^ (v >> 30u)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27u)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31u);
}

// Random draws for one pixel (reference SampleGenerator, sample_generator.hpp:31-61).
// Reference mode: u32 * 2^-32 as a double, one stream consumed by all the
// pixel's samples in order. Sample mode (the GPU's): each sample starts
// kSampleStride draws into the pixel's stream, and a draw is (u32 >> 8) * 2^-24,
// which FP32 holds exactly and never rounds to 1.
constexpr uint64_t kSampleStride = 1024;
struct Generator {
    Pcg32 pcg;
    bool float_draws{false};
    double next1D()
    {
        const uint32_t u = pcg.next();
        if (float_draws) return double(u >> 8u) * (1.0 / 16777216.0);
        return double(u) * (1.0 / 4294967296.0);
    }
    void next2D(double& x, double& y)
    {
        x = next1D();
        y = next1D();
    }
};
inline Generator pixelGenerator(uint64_t seed, uint32_t px, uint32_t py, uint32_t width)
{
    const uint64_t index = uint64_t(py) * uint64_t(width) + uint64_t(px);
    return Generator{Pcg32(seed, mixBits(index)), false};
}
inline Generator sampleGenerator(uint64_t seed, uint32_t px, uint32_t py, uint32_t width,
                                 uint32_t sample)
{
    Generator g = pixelGenerator(seed, px, py, width);
    g.float_draws = true;
    g.pcg.advance(uint64_t(sample) * kSampleStride);
    return g;
}

// Warps; reference sampling/warps.cpp:8-45.
inline void concentricDisk(double ux, double uy, double& dx, double& dy)
{
    const double ox = 2 * ux - 1, oy = 2 * uy - 1;
    if (ox == 0 && oy == 0) {
        dx = dy = 0;
        return;
    }
    double r, a;
    if (std::abs(ox) > std::abs(oy)) {
        r = ox;
        a = (kPi / 4) * (oy / ox);
    } else {
        r = oy;
        a = (kPi / 2) - (kPi / 4) * (ox / oy);
    }
    dx = r * std::cos(a);
    dy = r * std::sin(a);
}
inline V3d cosineHemisphere(double ux, double uy)
{
    double x, y;
    concentricDisk(ux, uy, x, y);
    return {x, y, std::sqrt(std::max(0.0, 1 - x * x - y * y))};
}
inline double cosineHemispherePdf(double c) { return c > 0 ? c * kInvPi : 0.0; }
inline V3d uniformSphere(double ux, double uy)
{
    const double z = 1 - 2 * ux;
    const double r = std::sqrt(std::max(0.0, 1 - z * z));
    const double phi = 2 * kPi * uy;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
inline double uniformSpherePdf() { return kInvPi / 4; }

// ---------------------------------------------------------------- geometry

// Sphere, reference geometry/sphere.cpp:66-135 (Haines discriminant, stable
// roots). Always FP64: the GPU evaluates the same FP64 expression, so the FP32
// mirror shares this function and rounds t afterwards.
inline bool sphereIntersect(V3d center, double radius, V3d o, V3d d, double tmin, double tmax,
                            double& t_out)
{
    const V3d oc = o - center;
    const double a = lengthSquared(d);
    if (!std::isnormal(a)) return false;
    const double h = dot(d, oc);
    const double c = lengthSquared(oc) - radius * radius;
    const V3d l = oc - d * (h / a);
    const double disc = a * (radius * radius - lengthSquared(l));
    if (disc < 0) return false;
    const double q = -(h + std::copysign(std::sqrt(disc), h));
    double near_t = 0, far_t = 0;
    if (q != 0) {
        near_t = q / a;
        far_t = c / q;
        if (near_t > far_t) std::swap(near_t, far_t);
    }
    double t = near_t;
    if (!(tmin < t && t < tmax)) {
        t = far_t;
        if (!(tmin < t && t < tmax)) return false;
    }
    const V3d p = o + d * t;
    if (!(lengthSquared(p - center) > 0)) return false;
    t_out = t;
    return true;
}

// Woop, Benthin and Wald 2013 watertight triangle test; reference
// geometry/triangle.cpp:95-176. For R = float the edge functions are recomputed
// in FP64 when one is exactly zero (Woop et al. sec. 3), as the GPU does.
template <class R>
inline R edgeFn(V3<R> a, V3<R> b) { return b.x * a.y - b.y * a.x; }

template <class R>
inline bool triangleIntersect(const V3<R> v[3], V3<R> o, V3<R> d, R tmin, R tmax, R& t_out,
                              R& b1_out, R& b2_out)
{
    if (!std::isnormal(lengthSquared(d))) return false;
    const R ax = std::abs(d.x), ay = std::abs(d.y), az = std::abs(d.z);
    int kz;
    if (ax > ay) kz = ax > az ? 0 : 2;
    else kz = ay > az ? 1 : 2;
    const R dz = d[kz];
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (dz < 0) std::swap(kx, ky);
    const R sx = d[kx] / dz, sy = d[ky] / dz, sz = R(1) / dz;
    auto xf = [&](V3<R> p) {
        const V3<R> off = p - o;
        const R z = off[kz];
        return V3<R>{off[kx] - sx * z, off[ky] - sy * z, sz * z};
    };
    const V3<R> a = xf(v[0]), b = xf(v[1]), c = xf(v[2]);
    R u = edgeFn(b, c), w1 = edgeFn(c, a), w2 = edgeFn(a, b);
    if constexpr (std::is_same_v<R, float>) {
        if (u == 0.0f || w1 == 0.0f || w2 == 0.0f) {
            u = float(double(c.x) * double(b.y) - double(c.y) * double(b.x));
            w1 = float(double(a.x) * double(c.y) - double(a.y) * double(c.x));
            w2 = float(double(b.x) * double(a.y) - double(b.y) * double(a.x));
        }
    }
    const bool inside = (u >= 0 && w1 >= 0 && w2 >= 0) || (u <= 0 && w1 <= 0 && w2 <= 0);
    if (!inside) return false;
    const R det = u + w1 + w2;
    if (det == 0) return false;
    const R t = (u * a.z + w1 * b.z + w2 * c.z) / det;
    if (!(tmin < t && t < tmax)) return false;
    t_out = t;
    b1_out = w1 / det;
    b2_out = w2 / det;
    return true;
}

// Reference Triangle::formsTriangle (triangle.cpp:198-205) on FP64 values.
inline bool formsTriangle(V3d p0, V3d p1, V3d p2)
{
    if (!isFiniteV(p0) || !isFiniteV(p1) || !isFiniteV(p2)) return false;
    return std::isnormal(lengthSquared(cross(p1 - p0, p2 - p0)));
}

// Row-major 3x4 affine map, applied as in core/transform.cpp:191-210.
template <class R>
struct Affine {
    R m[12];
    V3<R> point(V3<R> p) const
    {
        return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
                m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
    }
    V3<R> vector(V3<R> v) const
    {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
                m[8] * v.x + m[9] * v.y + m[10] * v.z};
    }
    // Transpose of the linear part times v: normals under the inverse map.
    V3<R> transposeVector(V3<R> v) const
    {
        return {m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
                m[2] * v.x + m[6] * v.y + m[10] * v.z};
    }
};

// ------------------------------------------------------------------- boxes

struct Box {
    V3d lo{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
           std::numeric_limits<double>::max()};
    V3d hi{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
           std::numeric_limits<double>::lowest()};
    bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
    void grow(V3d p)
    {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    void grow(const Box& b)
    {
        lo = {std::min(lo.x, b.lo.x), std::min(lo.y, b.lo.y), std::min(lo.z, b.lo.z)};
        hi = {std::max(hi.x, b.hi.x), std::max(hi.y, b.hi.y), std::max(hi.z, b.hi.z)};
    }
    double area() const
    {
        if (empty()) return 0;
        const V3d e = hi - lo;
        return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
};

// Arvo's box transform, reference core/transform.cpp:118-132, 212-225.
inline Box transformBox(const Affine<double>& a, const Box& b)
{
    Box out;
    double lo[3], hi[3];
    const double bl[3] = {b.lo.x, b.lo.y, b.lo.z}, bh[3] = {b.hi.x, b.hi.y, b.hi.z};
    for (int r = 0; r < 3; ++r) {
        double mn = 0, mx = 0;
        for (int k = 0; k < 3; ++k) {
            const double p = a.m[r * 4 + k] * bl[k], q = a.m[r * 4 + k] * bh[k];
            mn = k == 0 ? std::min(p, q) : mn + std::min(p, q);
            mx = k == 0 ? std::max(p, q) : mx + std::max(p, q);
        }
        lo[r] = mn + a.m[r * 4 + 3];
        hi[r] = mx + a.m[r * 4 + 3];
    }
    out.lo = {lo[0], lo[1], lo[2]};
    out.hi = {hi[0], hi[1], hi[2]};
    return out;
}

// ------------------------------------------------------------------- BVH

// Binned-SAH BVH2, reference accel/bvh.cpp:24-270 (12 bins, traversal cost
// 0.125, leaves of at most 4, SAH above depth 32 and median splits below).
struct BvhNode {
    Box box;
    uint32_t offset{0}; // leaf: first entry in order; interior: second child
    uint32_t count{0};  // 0 = interior
    uint8_t axis{0};
};

class Bvh {
public:
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> order;
    uint32_t max_depth{0};

    void build(const std::vector<Box>& boxes)
    {
        nodes.clear();
        order.clear();
        max_depth = 0;
        if (boxes.empty()) return;
        std::vector<Prim> prims(boxes.size());
        for (size_t i = 0; i < boxes.size(); ++i) {
            const Box& b = boxes[i];
            prims[i] = {b, {b.lo.x / 2 + b.hi.x / 2, b.lo.y / 2 + b.hi.y / 2, b.lo.z / 2 + b.hi.z / 2},
                        uint32_t(i)};
        }
        nodes.reserve(2 * prims.size());
        buildRec(prims.data(), prims.size(), 0, 0);
        order.resize(prims.size());
        for (size_t i = 0; i < prims.size(); ++i) order[i] = prims[i].index;
    }

    // Ordered depth-first traversal (accel/bvh.hpp:239-278). `range(lo, hi)`
    // supplies the current t range before each box test, so callers may shrink
    // it; `visit(prim)` returns true to stop.
    template <class RangeFn, class Visit>
    void traverse(V3d o, V3d d, RangeFn range, Visit visit) const
    {
        if (nodes.empty()) return;
        const double inv[3] = {1 / d.x, 1 / d.y, 1 / d.z};
        const double org[3] = {o.x, o.y, o.z};
        const unsigned neg = (d.x < 0 ? 1u : 0u) | (d.y < 0 ? 2u : 0u) | (d.z < 0 ? 4u : 0u);
        uint32_t stack[160];
        int sp = 0;
        uint32_t cur = 0;
        // 1 + 2 gamma(3) in double (Ize 2013), accel/bvh.hpp:84-85.
        const double eps = std::numeric_limits<double>::epsilon();
        const double widen = 1 + 2 * (3 * eps / 2) / (1 - 3 * eps / 2);
        while (true) {
            const BvhNode& n = nodes[cur];
            ++tl_counts.nodes;
            double lo, hi;
            range(lo, hi);
            const double bl[3] = {n.box.lo.x, n.box.lo.y, n.box.lo.z};
            const double bh[3] = {n.box.hi.x, n.box.hi.y, n.box.hi.z};
            for (int k = 0; k < 3; ++k) {
                double tn = (bl[k] - org[k]) * inv[k], tf = (bh[k] - org[k]) * inv[k];
                if (tn > tf) std::swap(tn, tf);
                tf *= tf < 0 ? 2 - widen : widen;
                lo = tn > lo ? tn : lo;
                hi = tf < hi ? tf : hi;
            }
            if (lo <= hi) {
                if (n.count == 0) {
                    uint32_t first = cur + 1, second = n.offset;
                    if ((neg >> n.axis) & 1u) std::swap(first, second);
                    stack[sp++] = second;
                    cur = first;
                    continue;
                }
                for (uint32_t e = n.offset; e < n.offset + n.count; ++e) {
                    ++tl_counts.prims;
                    if (visit(order[e])) return;
                }
            }
            if (sp == 0) return;
            cur = stack[--sp];
        }
    }

private:
    struct Prim {
        Box box;
        V3d centroid;
        uint32_t index;
    };
    static constexpr int kBins = 12;
    static constexpr double kTraversalCost = 0.125;
    static constexpr size_t kMaxLeaf = 4;
    static constexpr uint32_t kSahDepth = 32;

    static double axisOf(V3d p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }
    static double halfDistance(double lo, double hi) { return hi / 2 - lo / 2; }
    static int binOf(V3d c, const Box& cb, int axis)
    {
        const double lo = axisOf(cb.lo, axis);
        const double f = halfDistance(lo, axisOf(c, axis)) / halfDistance(lo, axisOf(cb.hi, axis));
        const double lim = f > 0 ? std::min(f, 1.0) : 0.0;
        return std::min(int(lim * kBins), kBins - 1);
    }

    uint32_t buildRec(Prim* p, size_t n, uint32_t first, uint32_t depth)
    {
        max_depth = std::max(max_depth, depth);
        const uint32_t index = uint32_t(nodes.size());
        Box bounds, cb;
        for (size_t i = 0; i < n; ++i) {
            bounds.grow(p[i].box);
            cb.grow(p[i].centroid);
        }
        nodes.push_back({bounds, 0, 0, 0});
        size_t left = 0;
        int axis = 0;
        if (n > 1) {
            bool decided = false;
            if (depth < kSahDepth) {
                double best_cost = std::numeric_limits<double>::infinity();
                int best_axis = -1, best_bin = 0;
                for (int a = 0; a < 3; ++a) {
                    if (!(halfDistance(axisOf(cb.lo, a), axisOf(cb.hi, a)) > 0)) continue;
                    Box bb[kBins];
                    size_t bc[kBins] = {};
                    for (size_t i = 0; i < n; ++i) {
                        const int b = binOf(p[i].centroid, cb, a);
                        bb[b].grow(p[i].box);
                        ++bc[b];
                    }
                    Box rb[kBins];
                    size_t rc[kBins] = {};
                    for (int b = kBins - 1; b > 0; --b) {
                        Box nb = b + 1 < kBins ? rb[b] : Box{};
                        size_t nc = b + 1 < kBins ? rc[b] : 0;
                        nb.grow(bb[b]);
                        rb[b - 1] = nb;
                        rc[b - 1] = nc + bc[b];
                    }
                    Box lb;
                    size_t lc = 0;
                    double axis_best = std::numeric_limits<double>::infinity();
                    int axis_bin = 0;
                    for (int b = 0; b + 1 < kBins; ++b) {
                        lb.grow(bb[b]);
                        lc += bc[b];
                        const double cost = lb.area() * double(lc) + rb[b].area() * double(rc[b]);
                        if (cost < axis_best) {
                            axis_best = cost;
                            axis_bin = b;
                        }
                    }
                    if (best_axis < 0 || axis_best < best_cost) {
                        best_cost = axis_best;
                        best_axis = a;
                        best_bin = axis_bin;
                    }
                }
                if (best_axis >= 0) {
                    const double area = bounds.area();
                    const bool leaf_cheaper = double(n) * area <= kTraversalCost * area + best_cost;
                    if (n <= kMaxLeaf && leaf_cheaper) {
                        left = 0;
                    } else {
                        Prim* mid = std::partition(p, p + n, [&](const Prim& q) {
                            return binOf(q.centroid, cb, best_axis) <= best_bin;
                        });
                        left = size_t(mid - p);
                        axis = best_axis;
                    }
                    decided = true;
                }
            }
            if (!decided) {
                if (n <= kMaxLeaf) {
                    left = 0;
                } else {
                    const V3d e = cb.hi - cb.lo;
                    axis = (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
                    left = n / 2;
                    std::nth_element(p, p + left, p + n, [&](const Prim& a, const Prim& b) {
                        return axisOf(a.centroid, axis) < axisOf(b.centroid, axis);
                    });
                }
            }
        }
        if (left == 0) {
            nodes[index].offset = first;
            nodes[index].count = uint32_t(n);
            return index;
        }
        buildRec(p, left, first, depth + 1);
        const uint32_t second = buildRec(p + left, n - left, first + uint32_t(left), depth + 1);
        nodes[index].offset = second;
        nodes[index].axis = uint8_t(axis);
        return index;
    }
};

} // namespace oracle
