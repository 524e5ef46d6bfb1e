// Per-element steps of the device BLAS build (blas_build.hip), written once for
// the GPU kernels and for host code: the CPU check runs the same steps in order
// over a mesh and verifies the tree it gets.
#pragma once

#include <cstdint>

#include "layout.hpp"

#if defined(__HIPCC__)
#define AB_HD __host__ __device__
#else
#define AB_HD
#endif

namespace ab {
namespace blas {

constexpr uint32_t kLeafMax = 4; // triangles per leaf (reference accel/bvh.cpp:33)

struct Box {
    float lo[3], hi[3];
};

// Leading zeros of x != 0 (keys are unique, so neighbours always differ).
AB_HD inline int clz64(unsigned long long x) { return __builtin_clzll(x); }

// The low 21 bits of x spread to every third bit.
AB_HD inline unsigned long long spread3(unsigned long long x)
{
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

AB_HD inline Box triBox(const DevTri& t)
{
    Box b;
    for (int k = 0; k < 3; ++k) {
        const float a = t.p0[k], c = t.p1[k], d = t.p2[k];
        b.lo[k] = a < c ? (a < d ? a : d) : (c < d ? c : d);
        b.hi[k] = a > c ? (a > d ? a : d) : (c > d ? c : d);
    }
    return b;
}

AB_HD inline float centroid(const Box& b, int k) { return 0.5f * b.lo[k] + 0.5f * b.hi[k]; }

// Morton code of the centroid quantised to `bits` per axis inside the centroid
// bounds [cmin, cmax], above the triangle's index i (idx_bits wide): unique keys.
AB_HD inline unsigned long long mortonKey(const Box& b, const float* cmin, const float* cmax, int bits, int idx_bits,
                                          uint32_t i)
{
    const float cells = float(1u << bits);
    unsigned long long code = 0;
    for (int k = 0; k < 3; ++k) {
        const float ext = cmax[k] - cmin[k];
        float q = ext > 0.0f ? (centroid(b, k) - cmin[k]) / ext * cells : 0.0f;
        q = q < 0.0f ? 0.0f : (q > cells - 1.0f ? cells - 1.0f : q);
        code |= spread3((unsigned long long)(q)) << k;
    }
    return (code << idx_bits) | i;
}

// The key the hierarchy is built over, for sorted position p: the top `top`
// bits of the Morton code, then p. Still sorted and unique; the tree splits on
// space for its top levels and halves position ranges below, which bounds its
// depth near log2(n) where Morton codes alone give long chains (uneven surfaces).
// Measured on the 1.3M-triangle displaced icosphere x 64 instances (2^22 rays):
// Morton alone 29 levels, 722 Mrays/s; 15 bits 25 levels, 836 Mrays/s; 12 bits
// 816; 6 bits 667; positions alone 20 levels, 475.
constexpr int kSplitTopBits = 15;
AB_HD inline unsigned long long splitKey(unsigned long long sorted_key, uint32_t p, int key_bits, int idx_bits,
                                         int top)
{
    const int morton_bits = key_bits - idx_bits;
    if (top > morton_bits) top = morton_bits;
    const unsigned long long morton = sorted_key >> idx_bits;
    const unsigned long long head = top > 0 ? morton >> (morton_bits - top) : 0ull;
    return (head << 32) | p;
}
static_assert(kSplitTopBits >= 0 && kSplitTopBits <= 32, "split key: head bits above a 32-bit position");

// Length of the common prefix of keys a and b; -1 when b is outside [0, n).
AB_HD inline int prefix(const unsigned long long* keys, int n, int a, int b)
{
    if (b < 0 || b >= n) return -1;
    return clz64(keys[a] ^ keys[b]);
}

// Internal node i of the n - 1 (Karras 2012, figure 4) over sorted unique keys.
// Children: >= 0 internal node, < 0 ~leaf (sorted position). Parent slots:
// leaves [0, n), internal node j at n + j; the root's parent is ~0.
AB_HD inline void hierarchyNode(const unsigned long long* keys, int n, int i, int32_t* child, uint32_t* parent,
                                uint32_t* first, uint32_t* count)
{
    if (i == 0) parent[n] = ~0u;
    const int d = prefix(keys, n, i, i + 1) - prefix(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = prefix(keys, n, i, i - d);
    int lmax = 2;
    while (prefix(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int32_t c0 = lo == gamma ? ~gamma : gamma;
    const int32_t c1 = hi == gamma + 1 ? ~(gamma + 1) : gamma + 1;
    child[2 * i] = c0;
    child[2 * i + 1] = c1;
    parent[c0 < 0 ? ~c0 : n + c0] = uint32_t(n + i);
    parent[c1 < 0 ? ~c1 : n + c1] = uint32_t(n + i);
    first[i] = uint32_t(lo);
    count[i] = uint32_t(hi - lo + 1);
}

// Box slot of a child: internal node j at n + j; a leaf (sorted position p) at
// its triangle's original index, where the triangle boxes were written.
AB_HD inline uint32_t boxSlot(int32_t c, uint32_t n, const unsigned long long* sorted, unsigned long long mask)
{
    return c < 0 ? uint32_t(sorted[~c] & mask) : n + uint32_t(c);
}

// Levels of the emitted subtree of internal node i from its children's (a leaf
// child counts one level; a node with at most kLeafMax triangles is a leaf).
AB_HD inline uint32_t nodeLevels(uint32_t count, int32_t c0, int32_t c1, uint32_t v0, uint32_t v1)
{
    if (count <= kLeafMax) return 1u;
    const uint32_t a = c0 < 0 ? 1u : v0, b = c1 < 0 ? 1u : v1;
    return 1u + (a > b ? a : b);
}

AB_HD inline void putBox(DevNode& nd, int slot, const Box& b)
{
    if (slot == 0) {
        nd.a[0] = b.lo[0]; nd.a[1] = b.hi[0]; nd.a[2] = b.lo[1]; nd.a[3] = b.hi[1]; nd.b[0] = b.lo[2]; nd.b[1] = b.hi[2];
    } else {
        nd.b[2] = b.lo[0]; nd.b[3] = b.hi[0]; nd.c[0] = b.lo[1]; nd.c[1] = b.hi[1]; nd.c[2] = b.lo[2]; nd.c[3] = b.hi[2];
    }
}

// The emitted node of kept internal node i. boxes: slot-indexed (see boxSlot).
AB_HD inline DevNode emitNode(uint32_t n, int i, const unsigned long long* sorted, unsigned long long mask,
                              const Box* boxes, const int32_t* child, const uint32_t* first, const uint32_t* count,
                              const uint32_t* keep, const uint32_t* index, uint32_t node_base, uint32_t tri_base)
{
    DevNode nd{};
    for (int c = 0; c < 2; ++c) {
        const int32_t ch = child[2 * i + c];
        int32_t ref;
        if (ch < 0) ref = encodeLeaf(tri_base + uint32_t(~ch), 1);
        else ref = keep[ch] ? int32_t(node_base + index[ch]) : encodeLeaf(tri_base + first[ch], count[ch]);
        putBox(nd, c, boxes[boxSlot(ch, n, sorted, mask)]);
        nd.child[c] = ref;
    }
    return nd;
}

// Bits per axis of the Morton code and of the index, for n triangles.
AB_HD inline void keyBits(uint32_t n, int& bits, int& idx_bits)
{
    idx_bits = 1;
    while (idx_bits < 32 && (1ull << idx_bits) < n) ++idx_bits;
    bits = (64 - idx_bits) / 3;
    if (bits > 21) bits = 21;
}

} // namespace blas
} // namespace ab
