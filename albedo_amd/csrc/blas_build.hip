// Device BLAS build: linear BVH (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012).
//
//   1. per triangle: box and centroid; the centroids' bounds by wave reductions
//      and one atomic per wave (order-preserving float bits);
//   2. 64-bit keys: the Morton code of the centroid quantised in those bounds,
//      above the triangle's index (keys are unique, as the construction needs);
//   3. radix sort of the keys (hipCUB);
//   4. every internal node i finds its key range and split from the common
//      prefix lengths of neighbouring keys, independently of the others;
//   5. boxes bottom-up: a thread climbs from each leaf and stops at the first
//      node its sibling's thread has not finished (one atomic per node, no
//      waiting), so each internal box is written once, after both children;
//   6. internal nodes with at most 4 triangles below become leaves (the
//      reference's leaf size, accel/bvh.cpp:33); the others are numbered by a
//      scan and written as two-child nodes, triangles in sorted order.
// Boxes are min/max of FP32 vertices, so they enclose their triangles exactly.
// The per-element steps are in blas_core.hpp, shared with the CPU check.

#include "blas_build.hpp"
#include "blas_core.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace ab {

namespace {

using blas::Box;

constexpr uint32_t kBlock = 256;

uint32_t blocksFor(uint64_t n) { return uint32_t((n + kBlock - 1) / kBlock); }

// Float bits mapped so unsigned order is float order (finite values).
__device__ __forceinline__ uint32_t orderedBits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float fromOrdered(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ float waveMin(float v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float waveMax(float v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ void kInitBounds(uint32_t* bounds)
{
    if (threadIdx.x < 6) bounds[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
}

__global__ __launch_bounds__(kBlock) void kTriBoxes(const DevTri* tris, uint32_t n, Box* boxes, uint32_t* bounds)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const Box b = blas::triBox(tris[i]);
        boxes[i] = b;
        for (int k = 0; k < 3; ++k) cmin[k] = cmax[k] = blas::centroid(b, k);
    }
    for (int k = 0; k < 3; ++k) {
        const float a = waveMin(cmin[k]), b = waveMax(cmax[k]);
        if (__lane_id() == 0 && a <= b) {
            atomicMin(bounds + k, orderedBits(a));
            atomicMax(bounds + 3 + k, orderedBits(b));
        }
    }
}

__global__ __launch_bounds__(kBlock) void kKeys(const Box* boxes, uint32_t n, const uint32_t* bounds, int bits,
                                                 int idx_bits, unsigned long long* keys)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float cmin[3], cmax[3];
    for (int k = 0; k < 3; ++k) {
        cmin[k] = fromOrdered(bounds[k]);
        cmax[k] = fromOrdered(bounds[3 + k]);
    }
    keys[i] = blas::mortonKey(boxes[i], cmin, cmax, bits, idx_bits, i);
}

__global__ __launch_bounds__(kBlock) void kSplitKeys(const unsigned long long* sorted, uint32_t n, int key_bits,
                                                      int idx_bits, unsigned long long* out)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) out[p] = blas::splitKey(sorted[p], p, key_bits, idx_bits, blas::kSplitTopBits);
}

__global__ __launch_bounds__(kBlock) void kHierarchy(const unsigned long long* keys, uint32_t n, int32_t* child,
                                                      uint32_t* parent, uint32_t* first, uint32_t* count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i + 1 < n) blas::hierarchyNode(keys, int(n), int(i), child, parent, first, count);
}

// Boxes of internal nodes are written by other threads of the same kernel:
// read them at device scope, not through this CU's cache.
__device__ __forceinline__ Box loadBox(const Box* p)
{
    const uint32_t* u = reinterpret_cast<const uint32_t*>(p);
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = __uint_as_float(__hip_atomic_load(u + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        b.hi[k] = __uint_as_float(__hip_atomic_load(u + 3 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    return b;
}

__global__ __launch_bounds__(kBlock) void kBottomUp(uint32_t n, int idx_bits, const unsigned long long* sorted,
                                                     Box* boxes, const int32_t* child, const uint32_t* parent,
                                                     const uint32_t* count, uint32_t* levels, uint32_t* visits)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const unsigned long long mask = (1ull << idx_bits) - 1ull;
    uint32_t node = parent[k];
    while (node != ~0u) {
        const uint32_t in = node - n;
        __threadfence();
        if (atomicAdd(visits + in, 1u) == 0) return; // the other child's thread goes on
        __threadfence();
        const int32_t c0 = child[2 * in], c1 = child[2 * in + 1];
        const Box b0 = loadBox(boxes + blas::boxSlot(c0, n, sorted, mask));
        const Box b1 = loadBox(boxes + blas::boxSlot(c1, n, sorted, mask));
        Box b;
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = fminf(b0.lo[a], b1.lo[a]);
            b.hi[a] = fmaxf(b0.hi[a], b1.hi[a]);
        }
        boxes[node] = b;
        const uint32_t v0 = c0 < 0 ? 1u : __hip_atomic_load(levels + c0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t v1 = c1 < 0 ? 1u : __hip_atomic_load(levels + c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        levels[in] = blas::nodeLevels(count[in], c0, c1, v0, v1);
        node = parent[node];
    }
}

__global__ __launch_bounds__(kBlock) void kKeep(const uint32_t* count, uint32_t m, uint32_t* keep)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) keep[i] = count[i] > blas::kLeafMax ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void kEmit(uint32_t n, int idx_bits, const unsigned long long* sorted,
                                                 const Box* boxes, const int32_t* child, const uint32_t* first,
                                                 const uint32_t* count, const uint32_t* keep, const uint32_t* index,
                                                 DevNode* out, uint32_t node_base, uint32_t tri_base)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i + 1 >= n || !keep[i]) return;
    const unsigned long long mask = (1ull << idx_bits) - 1ull;
    out[index[i]] = blas::emitNode(n, int(i), sorted, mask, boxes, child, first, count, keep, index, node_base,
                                   tri_base);
}

__global__ __launch_bounds__(kBlock) void kGather(const unsigned long long* sorted, uint32_t n, int idx_bits,
                                                   const DevTri* in, DevTri* out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = in[uint32_t(sorted[i] & ((1ull << idx_bits) - 1ull))];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

} // namespace

DeviceBlas::~DeviceBlas()
{
    if (mem_) (void)hipFree(mem_);
}

hipError_t DeviceBlas::prepare(const DevTri* host_tris, uint32_t n, hipStream_t s)
{
    if (n <= blas::kLeafMax || n >= (1u << 28) || mem_) return hipErrorInvalidValue;
    n_ = n;
    int bits = 0;
    blas::keyBits(n, bits, idx_bits_);
    const int key_bits = 3 * bits + idx_bits_;
    const size_t m = n - 1, slots = 2 * size_t(n) - 1;

    size_t sort_bytes = 0, scan_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (const unsigned long long*)nullptr,
                                                     (unsigned long long*)nullptr, int(n), 0, key_bits, s);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, int(m), s);
    if (e != hipSuccess) return e;
    scratch_bytes_ = std::max(sort_bytes, scan_bytes);

    struct Part {
        void** dst;
        size_t bytes;
    };
    Part parts[] = {{reinterpret_cast<void**>(&tris_), sizeof(DevTri) * n},
                    {reinterpret_cast<void**>(&boxes_), sizeof(Box) * slots},
                    {reinterpret_cast<void**>(&keys_), 8 * size_t(n)},
                    {reinterpret_cast<void**>(&sorted_), 8 * size_t(n)},
                    {reinterpret_cast<void**>(&child_), 8 * m},
                    {reinterpret_cast<void**>(&parent_), 4 * slots},
                    {reinterpret_cast<void**>(&first_), 4 * m},
                    {reinterpret_cast<void**>(&count_), 4 * m},
                    {reinterpret_cast<void**>(&levels_v_), 4 * m},
                    {reinterpret_cast<void**>(&visits_), 4 * m},
                    {reinterpret_cast<void**>(&keep_), 4 * m},
                    {reinterpret_cast<void**>(&index_), 4 * m},
                    {reinterpret_cast<void**>(&bounds_), 4 * 8},
                    {&scratch_, scratch_bytes_}};
    size_t total = 0;
    for (const Part& p : parts) total += align256(std::max<size_t>(p.bytes, 16));
    if ((e = hipMalloc(&mem_, total)) != hipSuccess) {
        mem_ = nullptr;
        return e;
    }
    size_t off = 0;
    for (const Part& p : parts) {
        *p.dst = static_cast<char*>(mem_) + off;
        off += align256(std::max<size_t>(p.bytes, 16));
    }
    Box* boxes = static_cast<Box*>(boxes_);

    if ((e = hipMemcpyAsync(tris_, host_tris, sizeof(DevTri) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(visits_, 0, 4 * m, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(kInitBounds, dim3(1), dim3(64), 0, s, bounds_);
    hipLaunchKernelGGL(kTriBoxes, dim3(blocksFor(n)), dim3(kBlock), 0, s, tris_, n, boxes, bounds_);
    hipLaunchKernelGGL(kKeys, dim3(blocksFor(n)), dim3(kBlock), 0, s, boxes, n, bounds_, bits, idx_bits_, keys_);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = scratch_bytes_;
    if ((e = hipcub::DeviceRadixSort::SortKeys(scratch_, bytes, keys_, sorted_, int(n), 0, key_bits, s)) != hipSuccess)
        return e;
    // The unsorted keys are spent: their buffer takes the split keys.
    hipLaunchKernelGGL(kSplitKeys, dim3(blocksFor(n)), dim3(kBlock), 0, s, sorted_, n, key_bits, idx_bits_, keys_);
    hipLaunchKernelGGL(kHierarchy, dim3(blocksFor(m)), dim3(kBlock), 0, s, keys_, n, child_, parent_, first_, count_);
    hipLaunchKernelGGL(kBottomUp, dim3(blocksFor(n)), dim3(kBlock), 0, s, n, idx_bits_, sorted_, boxes, child_,
                       parent_, count_, levels_v_, visits_);
    hipLaunchKernelGGL(kKeep, dim3(blocksFor(m)), dim3(kBlock), 0, s, count_, uint32_t(m), keep_);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    bytes = scratch_bytes_;
    if ((e = hipcub::DeviceScan::ExclusiveSum(scratch_, bytes, keep_, index_, int(m), s)) != hipSuccess) return e;
    uint32_t last_index = 0, last_keep = 0, root_levels = 0;
    if ((e = hipMemcpyAsync(&last_index, index_ + m - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&last_keep, keep_ + m - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&root_levels, levels_v_, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    nodes_ = last_index + last_keep;
    levels_ = root_levels;
    // More than 4 triangles: the root is a node, and the scan numbers it first.
    return nodes_ >= 1 && levels_ >= 2 ? hipSuccess : hipErrorUnknown;
}

hipError_t DeviceBlas::emit(DevNode* out_nodes, uint32_t node_base, DevTri* out_tris, uint32_t tri_base,
                            hipStream_t s)
{
    if (!mem_) return hipErrorInvalidValue;
    const uint32_t n = n_;
    hipLaunchKernelGGL(kEmit, dim3(blocksFor(n - 1)), dim3(kBlock), 0, s, n, idx_bits_, sorted_,
                       static_cast<const Box*>(boxes_), child_, first_, count_, keep_, index_, out_nodes, node_base,
                       tri_base);
    hipLaunchKernelGGL(kGather, dim3(blocksFor(n)), dim3(kBlock), 0, s, sorted_, n, idx_bits_, tris_, out_tris);
    return hipGetLastError();
}

} // namespace ab
