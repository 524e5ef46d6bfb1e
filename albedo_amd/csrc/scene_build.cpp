// Host scene build: validation (as the reference constructors validate), binned
// SAH BVH construction, and flattening into the 64-byte two-child node layout.
//
// The builder follows the reference's method and parameters (accel/bvh.cpp:24-32:
// 12 bins, leaves of at most 4, surface-area splits above depth 32 and median
// splits on the widest centroid axis below) except the traversal cost of triangle
// trees, priced for the GPU's query kernels (kCtTriangles). Trees need not match: closest hits are the
// lexicographic minimum of (t, object, prim), which no tree or traversal order
// changes.

#include "scene_build.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace ab {

namespace {

struct BBox {
    double lo[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double hi[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                    std::numeric_limits<double>::lowest()};
    void grow(const BBox& b)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const double* p)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
    double area() const
    {
        if (empty()) return 0;
        const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return 2 * (x * y + y * z + z * x);
    }
};

struct Item {
    BBox box;
    double c[3];
    uint32_t id;
};

struct TreeNode {
    BBox box;
    int32_t left{-1}, right{-1};
    uint32_t first{0}, count{0};
};

// Cost of a node visit relative to a primitive test, by tree. The reference's
// 0.125 (accel/bvh.cpp:24-32) prices an FP64 CPU. Over triangles, a node visit
// (both child boxes) costs about a triangle test in the query kernels here, and
// 1.0 measured fastest on MI355X (0.125, 0.5, 1 and 2 tried: incoherent trefoil
// rays 3 % faster, the full scene 4.5 % faster than 0.125). Over instances a
// "primitive" is a whole BLAS walk, so the TLAS keeps 0.125 (with 1.0 the 64-
// instance serving scene's queries were 7 % slower). Trees differ; hits do not.
constexpr double kCtTriangles = 1.0;
constexpr double kCtInstances = 0.125;

class SahBuilder {
public:
    explicit SahBuilder(double ct) : kCt(ct) {}
    std::vector<TreeNode> tree;
    uint32_t depth_reached{0};

    // Builds over `items` (reordered into leaf order); returns the root.
    int32_t build(std::vector<Item>& items)
    {
        tree.clear();
        depth_reached = 0;
        tree.reserve(2 * items.size());
        return node(items.data(), uint32_t(items.size()), 0, 0);
    }

private:
    const double kCt;
    static constexpr int kBins = 12;
    static constexpr uint32_t kMaxLeaf = 4;
    static constexpr uint32_t kSahDepth = 32;

    static double half(double lo, double hi) { return hi / 2 - lo / 2; }

    static int bin(const Item& it, const BBox& cb, int axis)
    {
        const double f = half(cb.lo[axis], it.c[axis]) / half(cb.lo[axis], cb.hi[axis]);
        const double lim = f > 0 ? std::min(f, 1.0) : 0.0;
        return std::min(int(lim * kBins), kBins - 1);
    }

    int32_t node(Item* it, uint32_t n, uint32_t first, uint32_t depth)
    {
        depth_reached = std::max(depth_reached, depth);
        BBox box, cb;
        for (uint32_t i = 0; i < n; ++i) {
            box.grow(it[i].box);
            cb.grow(it[i].c);
        }
        const int32_t self = int32_t(tree.size());
        tree.push_back({});
        tree[self].box = box;
        uint32_t left = 0;
        if (n > 1) {
            bool decided = false;
            if (depth < kSahDepth) {
                int best_axis = -1, best_bin = 0;
                double best_cost = std::numeric_limits<double>::infinity();
                for (int a = 0; a < 3; ++a) {
                    if (!(half(cb.lo[a], cb.hi[a]) > 0)) continue;
                    std::array<BBox, kBins> bb{};
                    std::array<uint32_t, kBins> bc{};
                    for (uint32_t i = 0; i < n; ++i) {
                        const int b = bin(it[i], cb, a);
                        bb[b].grow(it[i].box);
                        ++bc[b];
                    }
                    // Suffix sweep, then prefix sweep with the cost of each boundary.
                    std::array<double, kBins> right_cost{};
                    BBox acc;
                    uint32_t cnt = 0;
                    for (int b = kBins - 1; b > 0; --b) {
                        acc.grow(bb[b]);
                        cnt += bc[b];
                        right_cost[b - 1] = acc.area() * double(cnt);
                    }
                    BBox lacc;
                    uint32_t lcnt = 0;
                    for (int b = 0; b + 1 < kBins; ++b) {
                        lacc.grow(bb[b]);
                        lcnt += bc[b];
                        const double cost = lacc.area() * double(lcnt) + right_cost[b];
                        if (cost < best_cost) {
                            best_cost = cost;
                            best_axis = a;
                            best_bin = b;
                        }
                    }
                }
                if (best_axis >= 0) {
                    const double area = box.area();
                    const bool leaf_cheaper = double(n) * area <= kCt * area + best_cost;
                    if (!(n <= kMaxLeaf && leaf_cheaper)) {
                        Item* mid = std::stable_partition(
                            it, it + n, [&](const Item& x) { return bin(x, cb, best_axis) <= best_bin; });
                        left = uint32_t(mid - it);
                    }
                    decided = true;
                }
            }
            if (!decided && n > kMaxLeaf) {
                const double ex = cb.hi[0] - cb.lo[0], ey = cb.hi[1] - cb.lo[1], ez = cb.hi[2] - cb.lo[2];
                const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
                left = n / 2;
                std::nth_element(it, it + left, it + n,
                                 [&](const Item& x, const Item& y) { return x.c[axis] < y.c[axis]; });
            }
        }
        if (left == 0) {
            tree[self].first = first;
            tree[self].count = n;
            return self;
        }
        const int32_t l = node(it, left, first, depth + 1);
        const int32_t r = node(it + left, n - left, first + left, depth + 1);
        tree[self].left = l;
        tree[self].right = r;
        return self;
    }
};

void putBox(DevNode& n, int slot, const BBox& b)
{
    const float lo[3] = {float(b.lo[0]), float(b.lo[1]), float(b.lo[2])};
    const float hi[3] = {float(b.hi[0]), float(b.hi[1]), float(b.hi[2])};
    if (slot == 0) {
        n.a[0] = lo[0];
        n.a[1] = hi[0];
        n.a[2] = lo[1];
        n.a[3] = hi[1];
        n.b[0] = lo[2];
        n.b[1] = hi[2];
    } else {
        n.b[2] = lo[0];
        n.b[3] = hi[0];
        n.c[0] = lo[1];
        n.c[1] = hi[1];
        n.c[2] = lo[2];
        n.c[3] = hi[2];
    }
}

// Rounds a double box outward to float so the float box still encloses it.
BBox outward(const BBox& b)
{
    BBox r;
    for (int k = 0; k < 3; ++k) {
        float lo = float(b.lo[k]), hi = float(b.hi[k]);
        if (double(lo) > b.lo[k]) lo = std::nextafter(lo, -std::numeric_limits<float>::infinity());
        if (double(hi) < b.hi[k]) hi = std::nextafter(hi, std::numeric_limits<float>::infinity());
        r.lo[k] = lo;
        r.hi[k] = hi;
    }
    return r;
}

// Emits the device nodes of `tree` under `root` into `out`; leaves reference
// `leaf_base + first`. Returns the device index of the root node.
int32_t flatten(const SahBuilder& b, int32_t root, uint32_t leaf_base, std::vector<DevNode>& out)
{
    auto ref = [&](int32_t t, auto&& self) -> int32_t {
        const TreeNode& tn = b.tree[t];
        if (tn.left < 0) return encodeLeaf(leaf_base + tn.first, tn.count);
        const int32_t me = int32_t(out.size());
        out.push_back({});
        DevNode dn{};
        putBox(dn, 0, outward(b.tree[tn.left].box));
        putBox(dn, 1, outward(b.tree[tn.right].box));
        const int32_t l = self(tn.left, self);
        const int32_t r = self(tn.right, self);
        dn.child[0] = l;
        dn.child[1] = r;
        out[me] = dn;
        return me;
    };
    const TreeNode& rn = b.tree[root];
    if (rn.left < 0) {
        // A lone leaf: wrap it in a node whose second child is absent.
        const int32_t me = int32_t(out.size());
        DevNode dn{};
        putBox(dn, 0, outward(rn.box));
        BBox none;
        putBox(dn, 1, none);
        dn.child[0] = encodeLeaf(leaf_base + rn.first, rn.count);
        dn.child[1] = kNoChild;
        out.push_back(dn);
        return me;
    }
    return ref(root, ref);
}

// Emits the four-wide form of `tree` under `root` into `out`. A wide node is
// two consecutive DevNodes holding children 0-1 and 2-3, starting at an even
// index so it fills one 128-byte line; absent children are kNoChild. A node's
// children are gathered by opening, while it has fewer than four, the interior
// child of largest surface area (the usual collapse of a binary SAH tree;
// leaves stay as built). Returns the root reference; `stack` gets the most
// entries a traversal holds below this root: a visit pushes all but one of the
// children it enters, so the bound is the largest sum of (children - 1) over a
// root-to-node path.
int32_t flattenWide(const SahBuilder& b, int32_t root, uint32_t leaf_base, std::vector<DevNode>& out,
                    uint32_t& stack)
{
    auto emit = [&](int32_t t, uint32_t& bound, auto&& self) -> int32_t {
        const TreeNode& tn = b.tree[t];
        bound = 0;
        if (tn.left < 0) return encodeLeaf(leaf_base + tn.first, tn.count);
        int32_t kids[4] = {tn.left, tn.right, -1, -1};
        int nk = 2;
        while (nk < 4) {
            int open = -1;
            double open_area = -1;
            for (int i = 0; i < nk; ++i) {
                const TreeNode& c = b.tree[kids[i]];
                if (c.left >= 0 && c.box.area() > open_area) {
                    open = i;
                    open_area = c.box.area();
                }
            }
            if (open < 0) break;
            const TreeNode& c = b.tree[kids[open]];
            kids[open] = c.left;
            kids[nk++] = c.right;
        }
        if (out.size() & 1) out.push_back(DevNode{});
        const int32_t me = int32_t(out.size());
        out.resize(out.size() + 2);
        DevNode half[2]{};
        uint32_t below = 0;
        for (int i = 0; i < 4; ++i) {
            int32_t ref = kNoChild;
            if (i < nk) {
                uint32_t cb = 0;
                ref = self(kids[i], cb, self);
                below = std::max(below, cb);
                putBox(half[i >> 1], i & 1, outward(b.tree[kids[i]].box));
            } else {
                putBox(half[i >> 1], i & 1, BBox{});
            }
            half[i >> 1].child[i & 1] = ref;
        }
        out[me] = half[0];
        out[me + 1] = half[1];
        bound = below + uint32_t(nk - 1);
        return me;
    };
    const TreeNode& rn = b.tree[root];
    if (rn.left < 0) {
        // A lone leaf: a wide node with one child.
        if (out.size() & 1) out.push_back(DevNode{});
        const int32_t me = int32_t(out.size());
        DevNode half[2]{};
        putBox(half[0], 0, outward(rn.box));
        putBox(half[0], 1, BBox{});
        putBox(half[1], 0, BBox{});
        putBox(half[1], 1, BBox{});
        half[0].child[0] = encodeLeaf(leaf_base + rn.first, rn.count);
        half[0].child[1] = half[1].child[0] = half[1].child[1] = kNoChild;
        out.push_back(half[0]);
        out.push_back(half[1]);
        stack = 0;
        return me;
    }
    return emit(root, stack, emit);
}

bool finite3(const double* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Arvo's transformed box (reference core/transform.cpp:118-132).
BBox transformBox(const double* m, const BBox& b)
{
    BBox r;
    for (int row = 0; row < 3; ++row) {
        double mn = 0, mx = 0;
        for (int k = 0; k < 3; ++k) {
            const double p = m[row * 4 + k] * b.lo[k], q = m[row * 4 + k] * b.hi[k];
            mn = k == 0 ? std::min(p, q) : mn + std::min(p, q);
            mx = k == 0 ? std::max(p, q) : mx + std::max(p, q);
        }
        r.lo[row] = mn + m[row * 4 + 3];
        r.hi[row] = mx + m[row * 4 + 3];
    }
    return r;
}

struct MeshBuild {
    int32_t root{kNoChild};
    int32_t wide_root{kNoChild};
    uint32_t wide_stack{0};
    uint32_t shading_base{0};
    uint32_t tri_count{0}; // kept triangles
    uint32_t depth{0};
    uint32_t nodes{0};
    BBox box;
    bool smooth{false};
};

struct Hasher {
    uint64_t h = 0x243f6a8885a308d3ULL;
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            mix(w);
        }
        uint64_t tail = 0;
        std::memcpy(&tail, c + i, n - i);
        mix(tail ^ (uint64_t(n) << 56));
    }
    void mix(uint64_t w)
    {
        h ^= w + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
        h = (h ^ (h >> 31)) * 0xbf58476d1ce4e5b9ULL;
        h ^= h >> 29;
    }
};

} // namespace

uint64_t hashSceneDesc(const ab_scene_desc& d)
{
    Hasher hs;
    hs.mix(d.num_meshes);
    for (uint32_t i = 0; i < d.num_meshes; ++i) {
        const ab_mesh_desc& m = d.meshes[i];
        hs.mix(m.num_vertices);
        hs.mix(m.num_faces);
        hs.mix(uint64_t(m.smooth != 0) | (uint64_t(m.uv != nullptr) << 1));
        if (m.positions) hs.bytes(m.positions, sizeof(float) * 3 * m.num_vertices);
        if (m.indices) hs.bytes(m.indices, sizeof(uint32_t) * 3 * m.num_faces);
        if (m.uv) hs.bytes(m.uv, sizeof(float) * 6 * m.num_faces);
    }
    hs.mix(d.num_objects);
    if (d.num_objects) hs.bytes(d.objects, sizeof(ab_object_desc) * d.num_objects);
    hs.mix(d.num_materials);
    if (d.num_materials) hs.bytes(d.materials, sizeof(ab_material_desc) * d.num_materials);
    hs.mix(d.num_lights);
    if (d.num_lights) hs.bytes(d.lights, sizeof(ab_light_desc) * d.num_lights);
    hs.mix(d.build_flags);
    return hs.h;
}

ab_status buildHostScene(const ab_scene_desc& d, HostScene& out, std::string& err)
{
    out = HostScene{};
    if ((d.num_meshes && !d.meshes) || (d.num_objects && !d.objects) ||
        (d.num_materials && !d.materials) || (d.num_lights && !d.lights)) {
        err = "scene description has a null array with a non-zero count";
        return AB_INVALID_ARGUMENT;
    }
    if (d.build_flags & ~(AB_BUILD_DEVICE_BLAS | AB_BUILD_WIDE_BVH | AB_BUILD_ONE_PASS)) {
        err = "unknown scene build flags " + std::to_string(d.build_flags);
        return AB_INVALID_ARGUMENT;
    }
    if ((d.build_flags & AB_BUILD_DEVICE_BLAS) && (d.build_flags & AB_BUILD_WIDE_BVH)) {
        err = "AB_BUILD_WIDE_BVH covers host-built trees only; it cannot be combined with AB_BUILD_DEVICE_BLAS";
        return AB_INVALID_ARGUMENT;
    }
    const bool wide = (d.build_flags & AB_BUILD_WIDE_BVH) != 0;
    // ------------------------------------------------------------ materials
    for (uint32_t i = 0; i < d.num_materials; ++i) {
        const ab_material_desc& m = d.materials[i];
        DevMaterial dm{};
        dm.kind = m.kind;
        switch (m.kind) {
        case AB_MAT_DIFFUSE:
            if (m.albedo.kind > AB_TEX_PERLIN) {
                err = "material " + std::to_string(i) + " has an unknown texture kind";
                return AB_INVALID_ARGUMENT;
            }
            if ((m.albedo.kind == AB_TEX_CHECKER || m.albedo.kind == AB_TEX_PERLIN) &&
                !(m.albedo.frequency > 0 && std::isfinite(m.albedo.frequency))) {
                err = "material " + std::to_string(i) + ": checker frequency / perlin scale must be positive and finite";
                return AB_INVALID_ARGUMENT;
            }
            dm.tex_kind = m.albedo.kind;
            for (int k = 0; k < 3; ++k) {
                dm.a[k] = m.albedo.a[k];
                dm.b[k] = m.albedo.b[k];
            }
            dm.frequency = m.albedo.frequency;
            dm.veins = m.albedo.veins;
            break;
        case AB_MAT_CONDUCTOR:
            for (int k = 0; k < 3; ++k) {
                if (!(m.eta[k] > 0) || !(m.k[k] >= 0) || !std::isfinite(m.eta[k]) || !std::isfinite(m.k[k])) {
                    err = "material " + std::to_string(i) + ": conductor eta must be positive and k non-negative";
                    return AB_INVALID_ARGUMENT;
                }
                dm.a[k] = m.eta[k];
                dm.b[k] = m.k[k];
            }
            if (!(m.roughness >= 0) || !std::isfinite(m.roughness)) {
                err = "material " + std::to_string(i) + ": roughness must be finite and non-negative";
                return AB_INVALID_ARGUMENT;
            }
            dm.roughness = m.roughness;
            break;
        case AB_MAT_DIELECTRIC:
            if (!(m.ior > 0) || !std::isfinite(m.ior) || !(m.roughness >= 0) || !std::isfinite(m.roughness)) {
                err = "material " + std::to_string(i) + ": ior must be positive and roughness non-negative";
                return AB_INVALID_ARGUMENT;
            }
            dm.ior = m.ior;
            dm.roughness = m.roughness;
            break;
        case AB_MAT_MIX:
            if (m.first >= d.num_materials || m.second >= d.num_materials ||
                d.materials[m.first].kind == AB_MAT_MIX || d.materials[m.second].kind == AB_MAT_MIX) {
                err = "material " + std::to_string(i) + ": a mix needs two defined, non-mix materials";
                return AB_INVALID_ARGUMENT;
            }
            if (!(m.scale > 0) || !std::isfinite(m.scale) || !(m.veins >= 0) || !(m.threshold >= 0 && m.threshold <= 1)) {
                err = "material " + std::to_string(i) + ": mix scale must be positive, veins >= 0, threshold in [0, 1]";
                return AB_INVALID_ARGUMENT;
            }
            dm.first = m.first;
            dm.second = m.second;
            dm.mix_scale = m.scale;
            dm.mix_veins = m.veins;
            dm.mix_threshold = m.threshold;
            break;
        default:
            err = "material " + std::to_string(i) + " has an unknown kind";
            return AB_INVALID_ARGUMENT;
        }
        out.materials.push_back(dm);
    }

    // --------------------------------------------------------------- meshes
    std::vector<MeshBuild> mb(d.num_meshes);
    uint32_t blas_depth = 0;
    for (uint32_t mi = 0; mi < d.num_meshes; ++mi) {
        const ab_mesh_desc& m = d.meshes[mi];
        if (!m.positions || !m.indices) {
            err = "mesh " + std::to_string(mi) + " has no positions or faces";
            return AB_INVALID_ARGUMENT;
        }
        for (uint32_t v = 0; v < 3 * m.num_vertices; ++v)
            if (!std::isfinite(m.positions[v])) {
                err = "mesh position " + std::to_string(v / 3) + " has a non-finite coordinate";
                return AB_INVALID_ARGUMENT;
            }
        for (uint32_t f = 0; f < 3 * m.num_faces; ++f)
            if (m.indices[f] >= m.num_vertices) {
                err = "mesh face " + std::to_string(f / 3) + " uses position " + std::to_string(m.indices[f]) +
                      ", but the mesh has " + std::to_string(m.num_vertices) + " positions";
                return AB_INVALID_ARGUMENT;
            }
        auto P = [&](uint32_t i, double* p) {
            for (int k = 0; k < 3; ++k) p[k] = double(m.positions[3 * i + k]);
        };
        // Kept faces: those with a computable area (reference triangle_mesh.cpp:76-90).
        std::vector<uint32_t> kept;
        std::vector<std::array<double, 3>> vsum(m.smooth ? m.num_vertices : 0, {0, 0, 0});
        for (uint32_t f = 0; f < m.num_faces; ++f) {
            const uint32_t* ix = m.indices + 3 * f;
            double p0[3], p1[3], p2[3], e1[3], e2[3], an[3];
            P(ix[0], p0);
            P(ix[1], p1);
            P(ix[2], p2);
            for (int k = 0; k < 3; ++k) {
                e1[k] = p1[k] - p0[k];
                e2[k] = p2[k] - p0[k];
            }
            cross3(e1, e2, an);
            if (!std::isnormal(dot3(an, an))) continue;
            kept.push_back(f);
            if (m.smooth)
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k) vsum[ix[c]][k] += an[k];
        }
        if (kept.empty()) {
            err = "mesh geometry needs at least one triangle with area";
            return AB_INVALID_ARGUMENT;
        }
        MeshBuild& b = mb[mi];
        b.smooth = m.smooth != 0;
        b.shading_base = uint32_t(out.shading.size());
        b.tri_count = uint32_t(kept.size());
        const bool on_device = (d.build_flags & AB_BUILD_DEVICE_BLAS) && kept.size() > 4;
        // Unit vertex normals of a smooth mesh, once per vertex (none where the
        // sum has no length: those corners take the face normal).
        std::vector<std::array<float, 3>> vn(m.smooth ? m.num_vertices : 0);
        std::vector<uint8_t> vn_ok(m.smooth ? m.num_vertices : 0, 0);
        for (uint32_t v = 0; v < vn.size(); ++v) {
            const auto& s = vsum[v];
            const double l2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
            if (!std::isnormal(l2)) continue;
            const double l = std::sqrt(l2);
            for (int k = 0; k < 3; ++k) vn[v][k] = float(s[k] / l);
            vn_ok[v] = 1;
        }
        // Per triangle, independently: its box (the SAH input) and its shading
        // record (face normal, corner normals, uv); large meshes across threads.
        const uint32_t nt = uint32_t(kept.size());
        std::vector<Item> items(on_device ? 0 : nt);
        out.shading.resize(size_t(b.shading_base) + nt);
        const unsigned threads =
            nt < (1u << 16) ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<BBox> part_box(threads);
        auto work = [&](unsigned w) {
            const uint32_t t0 = uint32_t(uint64_t(nt) * w / threads), t1 = uint32_t(uint64_t(nt) * (w + 1) / threads);
            BBox mine;
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t* ix = m.indices + 3 * kept[t];
                double p[3][3];
                BBox tb;
                for (int c = 0; c < 3; ++c) {
                    P(ix[c], p[c]);
                    tb.grow(p[c]);
                }
                mine.grow(tb);
                if (!on_device) {
                    Item& it = items[t];
                    it.box = tb;
                    for (int k = 0; k < 3; ++k) it.c[k] = tb.lo[k] / 2 + tb.hi[k] / 2;
                    it.id = t;
                }
                DevTriShading sh{};
                double e1[3], e2[3], fn[3];
                for (int k = 0; k < 3; ++k) {
                    e1[k] = p[1][k] - p[0][k];
                    e2[k] = p[2][k] - p[0][k];
                }
                cross3(e1, e2, fn);
                for (int k = 0; k < 3; ++k) {
                    sh.p0[k] = m.positions[3 * ix[0] + k];
                    sh.p1[k] = m.positions[3 * ix[1] + k];
                    sh.p2[k] = m.positions[3 * ix[2] + k];
                }
                const double fl = std::sqrt(dot3(fn, fn));
                for (int k = 0; k < 3; ++k) sh.fn[k] = float(fn[k] / fl);
                float* cn[3] = {sh.n0, sh.n1, sh.n2};
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k) cn[c][k] = m.smooth && vn_ok[ix[c]] ? vn[ix[c]][k] : sh.fn[k];
                static const float kDefaultUv[6] = {0, 0, 1, 0, 1, 1};
                for (int k = 0; k < 6; ++k) sh.uv[k] = m.uv ? m.uv[6 * kept[t] + k] : kDefaultUv[k];
                out.shading[size_t(b.shading_base) + t] = sh;
            }
            part_box[w] = mine;
        };
        if (threads == 1) {
            work(0);
        } else {
            std::vector<std::thread> pool;
            for (unsigned w = 0; w < threads; ++w) pool.emplace_back(work, w);
            for (std::thread& th : pool) th.join();
        }
        for (const BBox& pb : part_box) b.box.grow(pb);
        if (on_device) {
            // Triangles in kept order; the device build sorts and places them.
            out.tris.reserve(out.tris.size() + nt);
            out.device_meshes.push_back({mi, uint32_t(out.tris.size()), uint32_t(kept.size())});
            for (uint32_t t = 0; t < kept.size(); ++t) {
                const uint32_t* ix = m.indices + 3 * kept[t];
                DevTri dt{};
                for (int k = 0; k < 3; ++k) {
                    dt.p0[k] = m.positions[3 * ix[0] + k];
                    dt.p1[k] = m.positions[3 * ix[1] + k];
                    dt.p2[k] = m.positions[3 * ix[2] + k];
                }
                dt.prim = t;
                out.tris.push_back(dt);
            }
            continue;
        }
        SahBuilder builder(kCtTriangles);
        const int32_t root = builder.build(items);
        const uint32_t tri_base = uint32_t(out.tris.size());
        for (const Item& it : items) {
            const uint32_t* ix = m.indices + 3 * kept[it.id];
            DevTri dt{};
            for (int k = 0; k < 3; ++k) {
                dt.p0[k] = m.positions[3 * ix[0] + k];
                dt.p1[k] = m.positions[3 * ix[1] + k];
                dt.p2[k] = m.positions[3 * ix[2] + k];
            }
            dt.prim = it.id;
            out.tris.push_back(dt);
        }
        const size_t before = out.nodes.size();
        b.root = flatten(builder, root, tri_base, out.nodes);
        b.nodes = uint32_t(out.nodes.size() - before);
        b.depth = builder.depth_reached + 1;
        out.blas_nodes += b.nodes;
        blas_depth = std::max(blas_depth, b.depth);
        if (wide) {
            b.wide_root = flattenWide(builder, root, tri_base, out.nodes, b.wide_stack);
            out.wide_blas_stack = std::max(out.wide_blas_stack, b.wide_stack);
        }
    }
    if (out.tris.size() >= (1u << (31 - kLeafCountBits))) {
        err = "scene has too many triangles for the leaf encoding";
        return AB_INVALID_ARGUMENT;
    }

    // -------------------------------------------------------------- objects
    std::vector<Item> items;
    for (uint32_t oi = 0; oi < d.num_objects; ++oi) {
        const ab_object_desc& o = d.objects[oi];
        if (o.material >= d.num_materials) {
            err = "object " + std::to_string(oi) + " uses an undefined material";
            return AB_INVALID_ARGUMENT;
        }
        if (o.light >= 0 && uint32_t(o.light) >= d.num_lights) {
            err = "object " + std::to_string(oi) + " names an undefined light";
            return AB_INVALID_ARGUMENT;
        }
        DevObject dobj{};
        dobj.kind = o.kind;
        dobj.material = o.material;
        dobj.light = o.light;
        dobj.blas_root = kNoChild;
        dobj.wide_root = kNoChild;
        Item it;
        it.id = oi;
        if (o.kind == AB_OBJ_SPHERE) {
            // reference geometry/sphere.cpp:21-41
            const double r = o.radius;
            if (!(r > 0) || !std::isfinite(r) || !std::isnormal(r * r)) {
                err = "sphere radius must be positive and its square a normal number";
                return AB_INVALID_ARGUMENT;
            }
            if (!finite3(o.center)) {
                err = "sphere center must be finite";
                return AB_INVALID_ARGUMENT;
            }
            for (int k = 0; k < 3; ++k) {
                it.box.lo[k] = o.center[k] - r;
                it.box.hi[k] = o.center[k] + r;
            }
            out.spheres.insert(out.spheres.end(), {o.center[0], o.center[1], o.center[2], r});
        } else if (o.kind == AB_OBJ_INSTANCE) {
            if (o.mesh >= d.num_meshes) {
                err = "instance " + std::to_string(oi) + " names an undefined mesh";
                return AB_INVALID_ARGUMENT;
            }
            for (int k = 0; k < 12; ++k)
                if (!std::isfinite(o.to_world[k]) || !std::isfinite(o.to_object[k])) {
                    err = "instance transform has a non-finite entry";
                    return AB_INVALID_ARGUMENT;
                }
            const MeshBuild& b = mb[o.mesh];
            it.box = transformBox(o.to_world, b.box);
            if (!finite3(it.box.lo) || !finite3(it.box.hi)) {
                err = "mesh instance transform takes the mesh out of range";
                return AB_INVALID_ARGUMENT;
            }
            dobj.blas_root = b.root;
            dobj.wide_root = b.wide_root;
            dobj.mesh = o.mesh;
            dobj.tri_shading_base = b.shading_base;
            dobj.tri_count = b.tri_count;
            dobj.smooth = b.smooth ? 1 : 0;
            for (int k = 0; k < 12; ++k) {
                dobj.to_object[k] = float(o.to_object[k]);
                dobj.to_world[k] = float(o.to_world[k]);
            }
            out.spheres.insert(out.spheres.end(), {0.0, 0.0, 0.0, 0.0});
        } else {
            err = "object " + std::to_string(oi) + " has an unknown kind";
            return AB_INVALID_ARGUMENT;
        }
        for (int k = 0; k < 3; ++k) it.c[k] = it.box.lo[k] / 2 + it.box.hi[k] / 2;
        items.push_back(it);
        out.objects.push_back(dobj);
    }

    // --------------------------------------------------------------- lights
    for (uint32_t li = 0; li < d.num_lights; ++li) {
        const ab_light_desc& l = d.lights[li];
        DevLight dl{};
        dl.kind = l.kind;
        switch (l.kind) {
        case AB_LIGHT_POINT:
            if (!finite3(l.position)) {
                err = "point light position must be finite";
                return AB_INVALID_ARGUMENT;
            }
            for (int k = 0; k < 3; ++k) {
                dl.position[k] = l.position[k];
                dl.a[k] = l.intensity[k];
            }
            break;
        case AB_LIGHT_AREA:
            if (l.object >= d.num_objects || d.objects[l.object].kind != AB_OBJ_SPHERE ||
                d.objects[l.object].light != int32_t(li)) {
                err = "area light " + std::to_string(li) + " must name a sphere object that links back to it";
                return AB_INVALID_ARGUMENT;
            }
            dl.object = l.object;
            for (int k = 0; k < 3; ++k) {
                dl.position[k] = d.objects[l.object].center[k];
                dl.a[k] = l.radiance[k];
            }
            dl.radius = d.objects[l.object].radius;
            break;
        case AB_LIGHT_SKY:
            if (out.sky_light >= 0) {
                err = "a scene has at most one sky";
                return AB_INVALID_ARGUMENT;
            }
            out.sky_light = int32_t(li);
            for (int k = 0; k < 3; ++k) {
                dl.a[k] = l.nadir[k];
                dl.b[k] = l.zenith[k];
            }
            break;
        default:
            err = "light " + std::to_string(li) + " has an unknown kind";
            return AB_INVALID_ARGUMENT;
        }
        out.lights.push_back(dl);
    }

    // ----------------------------------------------------------------- TLAS
    if (!items.empty()) {
        SahBuilder builder(kCtInstances);
        const int32_t root = builder.build(items);
        const uint32_t leaf_base = uint32_t(out.leafobj.size());
        for (const Item& it : items) out.leafobj.push_back(it.id);
        const size_t before = out.nodes.size();
        out.tlas_root = flatten(builder, root, leaf_base, out.nodes);
        out.tlas_nodes = out.nodes.size() - before;
        out.tlas_depth = builder.depth_reached + 1;
        if (wide) out.wide_tlas_root = flattenWide(builder, root, leaf_base, out.nodes, out.wide_tlas_stack);
        // The same objects split by kind, when both kinds are present (unless
        // the scene is to be traced in one pass). Moving instances of meshes of
        // up to 4 triangles (floor quads) to the sphere side was measured slower:
        // the full-feature scene 1.336 -> 1.366 s per image.
        std::vector<Item> spheres_only, instances_only;
        for (const Item& it : items)
            (out.objects[it.id].kind == kObjSphere ? spheres_only : instances_only).push_back(it);
        if (!spheres_only.empty() && !instances_only.empty() && !(d.build_flags & AB_BUILD_ONE_PASS)) {
            auto build_part = [&](std::vector<Item>& part, int32_t& part_root, uint32_t& part_depth) {
                SahBuilder b(kCtInstances);
                const int32_t r = b.build(part);
                const uint32_t base = uint32_t(out.leafobj.size());
                for (const Item& it : part) out.leafobj.push_back(it.id);
                part_root = flatten(b, r, base, out.nodes);
                part_depth = b.depth_reached + 1;
            };
            build_part(spheres_only, out.sphere_root, out.sphere_depth);
            build_part(instances_only, out.mesh_root, out.mesh_depth);
        }
    }
    out.max_blas_depth = blas_depth;
    out.hash = hashSceneDesc(d);
    return AB_OK;
}

} // namespace ab
