// Ray-query kernels: one lane per ray, LDS traversal stack, XCD-aware block
// order. Reference: Scene::intersect / Scene::occluded (scene/scene.cpp:107-165)
// over Bvh::nearestHit / anyHit (accel/bvh.hpp:211-267).

#include "kernels.hpp"
#include "traverse.hpp"

namespace ab {

namespace {

__device__ __forceinline__ uint32_t packBary(float b1, float b2)
{
    const float c1 = fminf(fmaxf(b1, 0.0f), 1.0f), c2 = fminf(fmaxf(b2, 0.0f), 1.0f);
    return uint32_t(c1 * 65535.0f + 0.5f) | (uint32_t(c2 * 65535.0f + 0.5f) << 16);
}

__device__ __forceinline__ RayF loadRay(const float4* rays, uint64_t i)
{
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    return RayF{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

template <int BLOCK, int STACK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(traversalWaves<STACK>(), 8))) void kClosest(DevScene sc, const float4* __restrict__ rays,
                                                  uint4* __restrict__ hits, uint64_t n)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const uint64_t i = uint64_t(xcdRemap(blockIdx.x, gridDim.x)) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const RayF r = loadRay(rays, i);
    HitF h;
    traceRay<BLOCK, false>(sc, r, h, LdsStack<BLOCK>{stack + threadIdx.x});
    const bool miss = h.object == kMissId;
    hits[i] = make_uint4(miss ? 0x7f800000u : __float_as_uint(h.t), h.object, h.prim,
                         miss ? 0u : packBary(h.b1, h.b2));
}

template <int BLOCK, int STACK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(traversalWaves<STACK>(), 8))) void kAny(DevScene sc, const float4* __restrict__ rays,
                                              uint8_t* __restrict__ occluded, uint64_t n)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const uint64_t i = uint64_t(xcdRemap(blockIdx.x, gridDim.x)) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const RayF r = loadRay(rays, i);
    HitF h;
    occluded[i] = traceRay<BLOCK, true>(sc, r, h, LdsStack<BLOCK>{stack + threadIdx.x}) ? 1 : 0;
}

template <int BLOCK, int STACK>
hipError_t launchC(const DevScene& sc, const ab_ray* rays, ab_hit* hits, uint64_t n, hipStream_t s)
{
    const uint64_t blocks = (n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((kClosest<BLOCK, STACK>), dim3(uint32_t(blocks)), dim3(BLOCK), 0, s, sc,
                       reinterpret_cast<const float4*>(rays), reinterpret_cast<uint4*>(hits), n);
    return hipGetLastError();
}

template <int BLOCK, int STACK>
hipError_t launchA(const DevScene& sc, const ab_ray* rays, uint8_t* occ, uint64_t n, hipStream_t s)
{
    const uint64_t blocks = (n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((kAny<BLOCK, STACK>), dim3(uint32_t(blocks)), dim3(BLOCK), 0, s, sc,
                       reinterpret_cast<const float4*>(rays), occ, n);
    return hipGetLastError();
}

} // namespace

hipError_t launchTraceClosest(const DevScene& sc, const ab_ray* rays, ab_hit* hits, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (n > (uint64_t(1) << 31) * 64) return hipErrorInvalidValue;
    if (sc.stack_depth <= kShortStack) return launchC<128, kShortStack>(sc, rays, hits, n, s);
    if (sc.stack_depth <= 32) return launchC<128, 32>(sc, rays, hits, n, s);
    if (sc.stack_depth <= 64) return launchC<128, 64>(sc, rays, hits, n, s);
    if (sc.stack_depth <= 128) return launchC<64, 128>(sc, rays, hits, n, s);
    if (sc.stack_depth <= kMaxStackDepth) return launchC<64, 256>(sc, rays, hits, n, s);
    return hipErrorInvalidValue;
}

hipError_t launchTraceAny(const DevScene& sc, const ab_ray* rays, uint8_t* occ, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (n > (uint64_t(1) << 31) * 64) return hipErrorInvalidValue;
    if (sc.stack_depth <= kShortStack) return launchA<128, kShortStack>(sc, rays, occ, n, s);
    if (sc.stack_depth <= 32) return launchA<128, 32>(sc, rays, occ, n, s);
    if (sc.stack_depth <= 64) return launchA<128, 64>(sc, rays, occ, n, s);
    if (sc.stack_depth <= 128) return launchA<64, 128>(sc, rays, occ, n, s);
    if (sc.stack_depth <= kMaxStackDepth) return launchA<64, 256>(sc, rays, occ, n, s);
    return hipErrorInvalidValue;
}

} // namespace ab
