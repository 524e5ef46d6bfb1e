// Runs the device BLAS build's per-element steps (albedo_amd/csrc/blas_core.hpp)
// on the CPU, in the order the kernels run them, over a mesh read from a file,
// and checks what they produce before any GPU run:
//   * keys are unique; the hierarchy references every leaf and every internal
//     node but the root exactly once, and children's ranges split the parent's;
//   * the bottom-up climb (one visit counter per node, a random leaf order)
//     visits every internal node twice and each box encloses its children;
//   * the emitted nodes: every reference in range, every triangle in exactly one
//     leaf, node boxes enclose their subtrees' triangles, the level count matches;
//   * closest hits through the emitted tree equal brute force on random rays.
// Usage: blas_check MESH.bin RAYS   (MESH.bin: u32 vertices, u32 faces, f32
// positions, u32 indices). Prints "ok n nodes levels" and exits 0 on success.
#include "blas_core.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <vector>

using namespace ab;
using blas::Box;

namespace {

int failures = 0;

void check(bool ok, const char* what, long long a = 0, long long b = 0)
{
    if (!ok && failures++ < 10) std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
}

bool encloses(const Box& outer, const Box& inner)
{
    for (int k = 0; k < 3; ++k)
        if (inner.lo[k] < outer.lo[k] || inner.hi[k] > outer.hi[k]) return false;
    return true;
}

// Double-precision ray-triangle test (Moller-Trumbore); the same function
// serves the tree walk and brute force, so only the tree can make them differ.
bool hitTri(const DevTri& t, const double* o, const double* d, double& tt)
{
    double e1[3], e2[3], p[3], q[3], s[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = double(t.p1[k]) - t.p0[k];
        e2[k] = double(t.p2[k]) - t.p0[k];
        s[k] = o[k] - t.p0[k];
    }
    p[0] = d[1] * e2[2] - d[2] * e2[1];
    p[1] = d[2] * e2[0] - d[0] * e2[2];
    p[2] = d[0] * e2[1] - d[1] * e2[0];
    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (std::fabs(det) < 1e-300) return false;
    const double inv = 1.0 / det;
    const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    if (u < 0 || u > 1) return false;
    q[0] = s[1] * e1[2] - s[2] * e1[1];
    q[1] = s[2] * e1[0] - s[0] * e1[2];
    q[2] = s[0] * e1[1] - s[1] * e1[0];
    const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0 || u + v > 1) return false;
    tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return tt > 0;
}

bool hitBox(const float* lo, const float* hi, const double* o, const double* d)
{
    double tn = 0, tf = std::numeric_limits<double>::infinity();
    for (int k = 0; k < 3; ++k) {
        // Widened by a relative 1e-9 so rounding never loses a box the ray touches.
        const double w = 1e-9 * (std::fabs(lo[k]) + std::fabs(hi[k]) + 1.0);
        const double a = (double(lo[k]) - w - o[k]) / d[k], b = (double(hi[k]) + w - o[k]) / d[k];
        tn = std::max(tn, std::min(a, b));
        tf = std::min(tf, std::max(a, b));
    }
    return tn <= tf;
}

Box childBox(const DevNode& n, int c)
{
    Box b;
    if (c == 0) {
        b.lo[0] = n.a[0]; b.hi[0] = n.a[1]; b.lo[1] = n.a[2]; b.hi[1] = n.a[3]; b.lo[2] = n.b[0]; b.hi[2] = n.b[1];
    } else {
        b.lo[0] = n.b[2]; b.hi[0] = n.b[3]; b.lo[1] = n.c[0]; b.hi[1] = n.c[1]; b.lo[2] = n.c[2]; b.hi[2] = n.c[3];
    }
    return b;
}

void decodeLeaf(int32_t ref, uint32_t& first, uint32_t& count)
{
    const uint32_t e = uint32_t(~ref);
    first = e >> kLeafCountBits;
    count = e & ((1u << kLeafCountBits) - 1);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: blas_check MESH.bin RAYS\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t nv = 0, nf = 0;
    if (std::fread(&nv, 4, 1, f) != 1 || std::fread(&nf, 4, 1, f) != 1) return 2;
    std::vector<float> pos(3 * size_t(nv));
    std::vector<uint32_t> idx(3 * size_t(nf));
    if (std::fread(pos.data(), 4, pos.size(), f) != pos.size() || std::fread(idx.data(), 4, idx.size(), f) != idx.size())
        return 2;
    std::fclose(f);
    std::vector<DevTri> tris(nf);
    for (uint32_t i = 0; i < nf; ++i) {
        DevTri& t = tris[i];
        for (int k = 0; k < 3; ++k) {
            t.p0[k] = pos[3 * idx[3 * i] + k];
            t.p1[k] = pos[3 * idx[3 * i + 1] + k];
            t.p2[k] = pos[3 * idx[3 * i + 2] + k];
        }
        t.prim = i;
        t.pad1 = t.pad2 = 0;
    }
    const uint32_t n = nf;
    if (n <= blas::kLeafMax) return 2;
    const uint32_t m = n - 1;

    // 1. Triangle boxes and centroid bounds (kTriBoxes).
    std::vector<Box> boxes(2 * size_t(n) - 1);
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        boxes[i] = blas::triBox(tris[i]);
        for (int k = 0; k < 3; ++k) {
            cmin[k] = std::min(cmin[k], blas::centroid(boxes[i], k));
            cmax[k] = std::max(cmax[k], blas::centroid(boxes[i], k));
        }
    }
    // 2-3. Keys and their sort (kKeys, radix sort).
    int bits = 0, idx_bits = 0;
    blas::keyBits(n, bits, idx_bits);
    check(3 * bits + idx_bits <= 64 && (1ull << idx_bits) >= n, "key bits", bits, idx_bits);
    std::vector<unsigned long long> sorted(n);
    for (uint32_t i = 0; i < n; ++i) sorted[i] = blas::mortonKey(boxes[i], cmin, cmax, bits, idx_bits, i);
    std::sort(sorted.begin(), sorted.end());
    for (uint32_t i = 1; i < n; ++i) check(sorted[i] != sorted[i - 1], "unique keys", i);
    const unsigned long long mask = (1ull << idx_bits) - 1ull;
    // The keys the hierarchy splits on (kSplitKeys): the Morton codes' top bits
    // and the sorted position (blas::kSplitTopBits unless given; a negative
    // value keeps the full sorted keys).
    const int top = argc > 3 ? std::atoi(argv[3]) : blas::kSplitTopBits;
    std::vector<unsigned long long> hkeys(sorted);
    if (top >= 0)
        for (uint32_t p = 0; p < n; ++p) hkeys[p] = blas::splitKey(sorted[p], p, 3 * bits + idx_bits, idx_bits, top);
    for (uint32_t i = 1; i < n; ++i) check(hkeys[i] > hkeys[i - 1], "split keys increase", i);

    // 4. Hierarchy (kHierarchy).
    std::vector<int32_t> child(2 * size_t(m), INT32_MIN);
    std::vector<uint32_t> parent(2 * size_t(n) - 1, 0xdeadbeefu), first(m), count(m);
    for (uint32_t i = 0; i < m; ++i)
        blas::hierarchyNode(hkeys.data(), int(n), int(i), child.data(), parent.data(), first.data(), count.data());
    std::vector<int> refs(2 * size_t(n) - 1, 0);
    for (uint32_t i = 0; i < m; ++i)
        for (int c = 0; c < 2; ++c) {
            const int32_t ch = child[2 * i + c];
            const bool in_range = ch < 0 ? uint32_t(~ch) < n : uint32_t(ch) < m && ch != 0;
            check(in_range, "child in range", i, ch);
            if (in_range) ++refs[ch < 0 ? ~ch : n + ch];
        }
    for (size_t s = 0; s < refs.size(); ++s)
        check(refs[s] == (s == n ? 0 : 1), "referenced once", static_cast<long long>(s), refs[s]);
    check(parent[n] == ~0u && first[0] == 0 && count[0] == n, "root covers all", first[0], count[0]);
    for (uint32_t i = 0; i < m; ++i) {
        uint32_t f0, c0, f1, c1;
        const int32_t a = child[2 * i], b = child[2 * i + 1];
        f0 = a < 0 ? uint32_t(~a) : first[a];
        c0 = a < 0 ? 1u : count[a];
        f1 = b < 0 ? uint32_t(~b) : first[b];
        c1 = b < 0 ? 1u : count[b];
        check(f0 == first[i] && f0 + c0 == f1 && f1 + c1 == first[i] + count[i], "children split the range", i, f0);
        check(parent[a < 0 ? ~a : n + a] == n + i && parent[b < 0 ? ~b : n + b] == n + i, "parent links", i);
    }
    if (failures) return 1;

    // 5. Bottom-up climb (kBottomUp), leaves in a random order.
    std::vector<uint32_t> visits(m, 0), levels(m, 0), order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::mt19937 rng(7);
    std::shuffle(order.begin(), order.end(), rng);
    for (uint32_t k : order) {
        uint32_t node = parent[k];
        while (node != ~0u) {
            const uint32_t in = node - n;
            if (visits[in]++ == 0) break;
            const int32_t c0 = child[2 * in], c1 = child[2 * in + 1];
            const Box& b0 = boxes[blas::boxSlot(c0, n, sorted.data(), mask)];
            const Box& b1 = boxes[blas::boxSlot(c1, n, sorted.data(), mask)];
            Box b;
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = std::min(b0.lo[a], b1.lo[a]);
                b.hi[a] = std::max(b0.hi[a], b1.hi[a]);
            }
            boxes[node] = b;
            levels[in] = blas::nodeLevels(count[in], c0, c1, c0 < 0 ? 1u : levels[c0], c1 < 0 ? 1u : levels[c1]);
            node = parent[node];
        }
    }
    for (uint32_t i = 0; i < m; ++i) {
        check(visits[i] == 2, "visited twice", i, visits[i]);
        check(encloses(boxes[n + i], boxes[blas::boxSlot(child[2 * i], n, sorted.data(), mask)]) &&
                  encloses(boxes[n + i], boxes[blas::boxSlot(child[2 * i + 1], n, sorted.data(), mask)]),
              "box encloses children", i);
    }
    if (failures) return 1;

    // 6. Keep, scan, emit, gather (kKeep, scan, kEmit, kGather).
    std::vector<uint32_t> keep(m), index(m);
    uint32_t total = 0;
    for (uint32_t i = 0; i < m; ++i) {
        keep[i] = count[i] > blas::kLeafMax ? 1u : 0u;
        index[i] = total;
        total += keep[i];
    }
    std::vector<DevNode> nodes(total);
    for (uint32_t i = 0; i < m; ++i)
        if (keep[i])
            nodes[index[i]] = blas::emitNode(n, int(i), sorted.data(), mask, boxes.data(), child.data(), first.data(),
                                             count.data(), keep.data(), index.data(), 0, 0);
    std::vector<DevTri> leaf_tris(n);
    for (uint32_t p = 0; p < n; ++p) leaf_tris[p] = tris[uint32_t(sorted[p] & mask)];
    check(keep[0] == 1 && index[0] == 0, "root emitted first");

    // Walk the emitted tree: references, coverage, enclosure, levels.
    std::vector<int> covered(n, 0);
    std::vector<int> reached(total, 0);
    uint32_t max_levels = 0;
    struct Item {
        int32_t ref;
        uint32_t level;
        Box box;
    };
    std::vector<Item> stack{{0, 1, boxes[n]}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        max_levels = std::max(max_levels, it.level);
        if (it.ref < 0) {
            uint32_t f0, c0;
            decodeLeaf(it.ref, f0, c0);
            check(c0 >= 1 && c0 <= blas::kLeafMax && f0 + c0 <= n, "leaf range", f0, c0);
            for (uint32_t t = f0; t < f0 + c0 && t < n; ++t) {
                ++covered[t];
                check(encloses(it.box, blas::triBox(leaf_tris[t])), "leaf box encloses triangle", t);
            }
            continue;
        }
        check(uint32_t(it.ref) < total, "node index", it.ref);
        if (uint32_t(it.ref) >= total) continue;
        ++reached[it.ref];
        const DevNode& nd = nodes[it.ref];
        for (int c = 0; c < 2; ++c) {
            const Box cb = childBox(nd, c);
            check(encloses(it.box, cb), "node box encloses child box", it.ref, c);
            stack.push_back({nd.child[c], it.level + 1, cb});
        }
    }
    for (uint32_t t = 0; t < n; ++t) check(covered[t] == 1, "triangle in one leaf", t, covered[t]);
    for (uint32_t i = 0; i < total; ++i) check(reached[i] == 1, "node reached once", i, reached[i]);
    check(max_levels == levels[0], "levels", max_levels, levels[0]);
    if (failures) return 1;

    // Closest hits through the tree equal brute force.
    const int rays = std::atoi(argv[2]);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = boxes[n].lo[k];
        hi[k] = boxes[n].hi[k];
    }
    int hits = 0;
    for (int r = 0; r < rays; ++r) {
        double o[3], d[3], target[3], len = 0;
        for (int k = 0; k < 3; ++k) {
            const double c = 0.5 * (lo[k] + hi[k]), e = hi[k] - lo[k] + 1e-3;
            o[k] = c + 1.5 * e * u(rng);
            target[k] = c + 0.5 * e * u(rng);
            d[k] = target[k] - o[k];
            len += d[k] * d[k];
        }
        for (int k = 0; k < 3; ++k) d[k] /= std::sqrt(len);
        double best_t = INFINITY, tt;
        uint32_t best = ~0u;
        for (uint32_t t = 0; t < n; ++t)
            if (hitTri(leaf_tris[t], o, d, tt) && (tt < best_t || (tt == best_t && leaf_tris[t].prim < best))) {
                best_t = tt;
                best = leaf_tris[t].prim;
            }
        double tree_t = INFINITY;
        uint32_t tree = ~0u;
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const int32_t ref = st.back();
            st.pop_back();
            if (ref < 0) {
                uint32_t f0, c0;
                decodeLeaf(ref, f0, c0);
                for (uint32_t t = f0; t < f0 + c0; ++t)
                    if (hitTri(leaf_tris[t], o, d, tt) &&
                        (tt < tree_t || (tt == tree_t && leaf_tris[t].prim < tree))) {
                        tree_t = tt;
                        tree = leaf_tris[t].prim;
                    }
                continue;
            }
            const DevNode& nd = nodes[ref];
            for (int c = 0; c < 2; ++c) {
                const Box cb = childBox(nd, c);
                if (hitBox(cb.lo, cb.hi, o, d)) st.push_back(nd.child[c]);
            }
        }
        check(tree == best && (best == ~0u || tree_t == best_t), "tree hit = brute force", r, int(tree));
        hits += best != ~0u;
    }
    if (failures) return 1;
    std::printf("ok n %u nodes %u levels %u rays %d hits %d\n", n, total, levels[0], rays, hits);
    return 0;
}
