// Ray-query kernels. Reference: Scene::intersect / Scene::occluded
// (scene/scene.cpp:107-165) over Bvh::nearestHit / anyHit (accel/bvh.hpp:211-267).
//
// Each wave owns a contiguous slice of rays and keeps its 64 lanes busy: a lane
// whose ray has finished takes the next ray of the slice (once at least kRefill
// lanes are free), so a wave no longer waits for its longest ray. Incoherent
// rays vary widely in traversal length; with one ray per lane the waves ran
// with about 16 % of their lanes active. Results are written per ray and do not
// depend on which lane or wave traced it. Slices are laid out so each XCD works
// on a contiguous range of rays.

#include "kernels.hpp"
#include "traverse.hpp"

#include <algorithm>

namespace ab {

namespace {

constexpr uint32_t kBlock = 64;       // one wave per block: waves retire independently
constexpr uint32_t kRefill = 16;      // free lanes that trigger a refill
constexpr uint32_t kSteps = 16;       // traversal steps between refill checks
// Waves a launch aims for: slices of 6 rays per lane for 2^24 rays. Measured on
// the trefoil sets: 8 per lane (32768 waves) 1.94 + 3.31 ms, 6 per lane
// 1.64 + 3.36 ms, 4 per lane 1.75 + 3.45 ms (coherent + incoherent).
constexpr uint32_t kTargetWaves = 43691;

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

// Registers the compiler may use: with the FP64 sphere test compiled out the
// query fits 72 VGPRs (7 waves per SIMD) with few spills; with it, 80 (6 waves).
// The four-wide visit needs 76 (6 waves).
template <bool SPHERES, bool WIDE>
constexpr int queryWaves() { return SPHERES || WIDE ? 6 : 7; }

// A stack class for four-wide trees: 26 entries (6.5 KB per wave) leave LDS
// room for 6 waves per SIMD; the trefoil's bound is 25.
constexpr uint32_t kWideStack = 26;

// One kernel for closest and any hit: `per_lane` rays per lane of each wave's slice.
template <int STACK, bool ANY, bool SPHERES, bool LEAF_IN_REG, bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(queryWaves<SPHERES, WIDE>(), 8))) void kQuery(
    DevScene sc, const float4* __restrict__ rays, uint4* __restrict__ hits, uint8_t* __restrict__ occluded, uint64_t n,
    uint32_t per_lane)
{
    __shared__ int32_t stack[STACK * kBlock];
    const uint32_t lane = __lane_id();
    const uint64_t slice = uint64_t(per_lane) * 64;
    uint64_t next = uint64_t(xcdRemap(blockIdx.x, gridDim.x)) * slice;
    const uint64_t end = next + slice < n ? next + slice : n;
    const LdsStack<kBlock> stk{stack + threadIdx.x};
    RayQuery<kBlock, ANY, SPHERES, LEAF_IN_REG, WIDE> q;
    constexpr uint64_t kNone = ~0ull;
    uint64_t id = kNone;
    while (true) {
        const bool idle = id == kNone;
        const unsigned long long free_lanes = __ballot(idle);
        if (next < end && (__popcll(free_lanes) >= kRefill || free_lanes == ~0ull)) {
            const uint64_t mine = next + uint64_t(__popcll(free_lanes & ((1ull << lane) - 1ull)));
            next += uint64_t(__popcll(free_lanes));
            if (idle && mine < end) {
                q.begin(sc, loadRay(rays, mine));
                id = mine;
            }
        }
        if (__ballot(id != kNone) == 0) break;
        if (id != kNone) {
            for (uint32_t k = 0; k < kSteps && !q.done; ++k) q.step(sc, stk);
            if (q.done) {
                if (ANY) {
                    occluded[id] = q.best.object != kMissId ? 1 : 0;
                } else {
                    const HitF h = q.result();
                    const bool miss = h.object == kMissId;
                    hits[id] = make_uint4(miss ? 0x7f800000u : __float_as_uint(h.t), h.object, h.prim,
                                          miss ? 0u : packBary(h.b1, h.b2));
                }
                id = kNone;
            }
        }
    }
}

template <int STACK, bool ANY, bool LEAF_IN_REG, bool WIDE>
hipError_t launchQ(const DevScene& sc, const ab_ray* rays, ab_hit* hits, uint8_t* occ, uint64_t n, hipStream_t s)
{
    const uint64_t per_lane = (n + 64ull * kTargetWaves - 1) / (64ull * kTargetWaves);
    const uint64_t waves = (n + 64 * per_lane - 1) / (64 * per_lane);
    const float4* r = reinterpret_cast<const float4*>(rays);
    uint4* h = reinterpret_cast<uint4*>(hits);
    if (sc.has_spheres)
        hipLaunchKernelGGL((kQuery<STACK, ANY, true, LEAF_IN_REG, WIDE>), dim3(uint32_t(waves)), dim3(kBlock), 0, s,
                           sc, r, h, occ, n, uint32_t(per_lane));
    else
        hipLaunchKernelGGL((kQuery<STACK, ANY, false, LEAF_IN_REG, WIDE>), dim3(uint32_t(waves)), dim3(kBlock), 0, s,
                           sc, r, h, occ, n, uint32_t(per_lane));
    return hipGetLastError();
}

// The query keeps the rest of a TLAS leaf on the stack when that extra entry
// stays within the scene's stack class (cheaper), else in a register.
template <int STACK, bool ANY, bool WIDE>
hipError_t launchClass(const DevScene& sc, uint32_t depth, const ab_ray* rays, ab_hit* hits, uint8_t* occ,
                       uint64_t n, hipStream_t s)
{
    if (depth + 1 <= uint32_t(STACK)) return launchQ<STACK, ANY, false, WIDE>(sc, rays, hits, occ, n, s);
    return launchQ<STACK, ANY, true, WIDE>(sc, rays, hits, occ, n, s);
}

template <bool ANY, bool WIDE>
hipError_t launchDepth(const DevScene& sc, uint32_t depth, const ab_ray* rays, ab_hit* hits, uint8_t* occ,
                       uint64_t n, hipStream_t s)
{
    if (depth <= kShortStack) return launchClass<kShortStack, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    if constexpr (WIDE) {
        if (depth + 1 <= kWideStack) return launchClass<kWideStack, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    }
    if (depth <= 32) return launchClass<32, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    if (depth <= 64) return launchClass<64, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    if (depth <= 128) return launchClass<128, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    if (depth <= kMaxStackDepth) return launchClass<256, ANY, WIDE>(sc, depth, rays, hits, occ, n, s);
    return hipErrorInvalidValue;
}

template <bool ANY>
hipError_t launchByStack(const DevScene& sc, const ab_ray* rays, ab_hit* hits, uint8_t* occ, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (n > (uint64_t(1) << 31) * 64) return hipErrorInvalidValue;
    if (sc.wide_root != kNoChild) return launchDepth<ANY, true>(sc, sc.wide_stack, rays, hits, occ, n, s);
    return launchDepth<ANY, false>(sc, sc.stack_depth, rays, hits, occ, n, s);
}

// Copy c of a triangle: its vertices' coordinates in the order (c + 1, c + 2, c)
// mod 3, the kept-face index kept beside the first vertex.
__global__ void kRotateTris(DevTri* tris, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const DevTri t = tris[2 * size_t(n) + i];
        DevTri a = t, b = t;
        const float* src[3] = {t.p0, t.p1, t.p2};
        float* da[3] = {a.p0, a.p1, a.p2};
        float* db[3] = {b.p0, b.p1, b.p2};
        for (int v = 0; v < 3; ++v) {
            da[v][0] = src[v][1]; // copy 0: (y, z, x)
            da[v][1] = src[v][2];
            da[v][2] = src[v][0];
            db[v][0] = src[v][2]; // copy 1: (z, x, y)
            db[v][1] = src[v][0];
            db[v][2] = src[v][1];
        }
        tris[i] = a;
        tris[size_t(n) + i] = b;
    }
}

} // namespace

hipError_t launchRotateTris(DevTri* tris, uint32_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(kRotateTris, dim3(blocks), dim3(256), 0, s, tris, n);
    return hipGetLastError();
}

hipError_t launchTraceClosest(const DevScene& sc, const ab_ray* rays, ab_hit* hits, uint64_t n, hipStream_t s)
{
    return launchByStack<false>(sc, rays, hits, nullptr, n, s);
}

hipError_t launchTraceAny(const DevScene& sc, const ab_ray* rays, uint8_t* occ, uint64_t n, hipStream_t s)
{
    return launchByStack<true>(sc, rays, nullptr, occ, n, s);
}

} // namespace ab
