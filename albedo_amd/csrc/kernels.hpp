// Host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>

#include "../../include/albedo_gpu.h"
#include "layout.hpp"

namespace ab {

// Closest-hit / any-hit over n rays; `stack_depth` is DevScene::stack_depth.
hipError_t launchTraceClosest(const DevScene& sc, const ab_ray* d_rays, ab_hit* d_hits, uint64_t n,
                              hipStream_t stream);
hipError_t launchTraceAny(const DevScene& sc, const ab_ray* d_rays, uint8_t* d_occluded, uint64_t n,
                          hipStream_t stream);

// Fills copies 0 and 1 of a scene's triangles (DevScene::tris) from copy 2,
// which holds them as built, coordinates in (x, y, z) order; n per copy.
hipError_t launchRotateTris(DevTri* tris, uint32_t n, hipStream_t stream);

// Largest stack depth the trace kernels support.
constexpr uint32_t kMaxStackDepth = 256;

// Traversal kernels keep their stack in LDS, STACK entries of 4 B per lane. With
// a 20-entry stack a wave needs 5 KiB, so LDS leaves room for 8 waves per SIMD
// and registers become the limit (the ray-query kernels then ask the compiler
// for 6-7 waves per SIMD); deeper stacks are LDS-bound.
constexpr uint32_t kShortStack = 20;
// A class between 20 and 32 for the path tracer's refill kernels: 6 KB of LDS per
// wave leaves room for 6 waves per SIMD where 32 entries allow 5 (the
// full-feature scene's bound is 22, one more with a TLAS leaf's rest on the stack).
constexpr uint32_t kRefillStack24 = 24;

// Maps block k of a G-block launch so each of the 8 XCDs (which receive blocks
// round-robin) works on one contiguous range of logical blocks, keeping nearby
// rays in one L2.
__host__ __device__ inline uint32_t xcdRemap(uint32_t k, uint32_t g)
{
    const uint32_t xcd = k & 7u, idx = k >> 3;
    const uint32_t q = g >> 3, r = g & 7u;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

} // namespace ab
