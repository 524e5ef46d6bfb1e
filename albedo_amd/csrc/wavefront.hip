// Wavefront path tracer kernels (reference estimator/path_estimator.cpp:73-185
// and render/render_image.cpp:27-44, rebuilt as generate -> extend -> bin ->
// shade -> shadow -> finish over queues in HBM).
//
// Every path is identified by (job, pixel, sample) and draws from its own PCG32
// stream position, so its result does not depend on which slot, wave or launch
// handles it. Radiance is accumulated per pixel as 2^-28 fixed point in 64-bit
// integers, whose sums do not depend on the order of the atomic adds: images are
// bit-identical for every launch configuration.

#include "kernels.hpp"
#include "shade.hpp"
#include "traverse.hpp"
#include "wavefront.hpp"

namespace ab {

namespace {

constexpr double kFixedScale = 268435456.0; // 2^28
static_assert(kFixedScale == double(1ull << AB_SUMS_FRACTION_BITS), "sums fixed point = the wire format's");
constexpr float kMaxSampleRadiance = 1e5f;  // per-sample clamp that keeps 64-bit sums exact

// The live segment holding global work index w: the last of [lo, hi) (logical
// ring indices, hi > lo) whose begin is <= w. Segment lo begins at or before
// every index still to be generated.
__device__ __forceinline__ const DevSegment& findSegment(const DevSegment* segs, uint32_t mask, uint32_t lo,
                                                         uint32_t hi, unsigned long long w)
{
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (segs[mid & mask].begin <= w) lo = mid;
        else hi = mid;
    }
    return segs[lo & mask];
}

__device__ __forceinline__ float powerHeuristic(float chosen, float other)
{
    const float c2 = chosen * chosen;
    const float t = c2 + other * other;
    return t > 0.0f ? c2 / t : 0.0f;
}

// The generation a bounce starts with, from the counters as the previous bounce
// left them: new samples first take the n_recycle recycled slots and positions
// (recycling is chosen only when the work left covers them, and work only
// grows), then free slots, appended to the queue, up to the work exposed. The
// slots the last finish freed (n_term - n_recycle, pushed above n_free) are free
// again. kGenerate and the extension kernels compute it; kPlanBins applies it to
// the counters after them.
struct Generation {
    uint32_t free_top;  // free slots
    uint32_t n_new;     // samples generated
    uint32_t n_recycle; // ... of which in recycled slots and positions
    uint32_t n_extra;   // ... of which in free slots, appended to the queue
};

__device__ __forceinline__ Generation pendingGeneration(const Counters* c)
{
    const uint32_t free_top = c->n_free + (c->n_term - c->n_recycle);
    const unsigned long long rest = c->work_total - c->work_next - c->n_recycle;
    const uint32_t extra = uint32_t(rest < free_top ? rest : free_top);
    return {free_top, c->n_recycle + extra, c->n_recycle, extra};
}

__device__ __forceinline__ uint32_t gridStart() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gridStride() { return gridDim.x * blockDim.x; }

// The range [lo, hi) of an order over n items that this block walks when the
// grid splits it into contiguous ranges (whole chunks of BLOCK), so that what
// the blocks write in turn keeps that order.
template <int BLOCK>
__device__ __forceinline__ void blockSpan(uint32_t n, uint32_t& lo, uint32_t& hi)
{
    const uint32_t per = ((n + gridDim.x - 1) / gridDim.x + BLOCK - 1) / BLOCK * BLOCK;
    lo = min(n, blockIdx.x * per);
    hi = min(n, lo + per);
}

// ---------------------------------------------------------------- generate

// Camera rays for the next n_new samples (reference render_image.cpp:27-44 and
// PerspectiveCamera::generateRay, perspective_camera.cpp:74-87). A job's samples
// are taken in passes of `group` samples per pixel: pass g covers samples
// [g G, g G + G) of every pixel (the last pass fewer), pixel-major inside the
// pass. Samples of one pixel then take neighbouring queue positions (adjacent
// when appended, the next recycled positions otherwise), so kFinish can often sum
// them within a wave before touching the pixel.
__global__ void kGenerate(PathState ps)
{
    const Counters* c = ps.ctr;
    const Generation g = pendingGeneration(c);
    const uint32_t n_new = g.n_new, n_rec = g.n_recycle, free_top = g.free_top, act0 = c->n_active;
    const unsigned long long w0 = c->work_next;
    const uint32_t seg_lo = c->seg_lo, seg_hi = c->seg_hi;
    for (uint32_t i = gridStart(); i < n_new; i += gridStride()) {
        // A recycled terminated path's position and slot, or a free slot
        // appended to the queue.
        uint32_t pos, slot;
        if (i < n_rec) {
            pos = ps.termq[i];
            slot = ps.active[pos] & kSlotMask;
        } else {
            pos = act0 + (i - n_rec);
            slot = ps.freeq[free_top - 1 - (i - n_rec)];
            ps.slot_free[slot] = 0;
        }
        const unsigned long long w = w0 + i;
        const DevSegment& sg = findSegment(ps.segs, ps.seg_mask, seg_lo, seg_hi, w);
        const uint32_t j = sg.job;
        const DevJob& job = ps.jobs[j];
        const unsigned long long local = sg.local + (w - sg.begin);
        const unsigned long long npx = uint64_t(job.width) * job.height;
        const uint32_t group = job.group, spp = job.spa * job.spa;
        const unsigned long long span = npx * group;
        const uint32_t pass = uint32_t(local / span);
        const unsigned long long r = local - uint64_t(pass) * span;
        const uint32_t first = pass * group;
        const uint32_t in_pass = min(group, spp - first);
        const uint32_t pix = uint32_t(r / in_pass);
        const uint32_t s = first + uint32_t(r - uint64_t(pix) * in_pass);
        const uint32_t x = pix % job.width, y = pix / job.width;
        Rng rng = sampleRng(job.seed, pix, s);
        const uint32_t sx = s % job.spa, sy = s / job.spa;
        const float jx = rng.uniform(), jy = rng.uniform();
        const float n = float(job.spa);
        const float fx = (float(x) + (float(sx) + jx) / n) / float(job.width);
        const float fy = (float(y) + (float(sy) + jy) / n) / float(job.height);
        const float h = 2.0f * fx - 1.0f, v = 1.0f - 2.0f * fy;
        const F3 fw = f3(job.cam_forward[0], job.cam_forward[1], job.cam_forward[2]);
        const F3 hw = f3(job.cam_hw[0], job.cam_hw[1], job.cam_hw[2]);
        const F3 hh = f3(job.cam_hh[0], job.cam_hh[1], job.cam_hh[2]);
        const F3 d = normalize(fw + hw * h + hh * v);
        const float ox = job.cam_origin[0], oy = job.cam_origin[1], oz = job.cam_origin[2];
        ps.org[slot] = make_float4(ox, oy, oz, 0.0f);
        ps.dir[slot] = make_float4(d.x, d.y, d.z, 1.0f);
        ps.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(0));
        ps.rad[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps.prev[slot] = make_float4(ox, oy, oz, __uint_as_float(j));
        ps.rng[slot] = rng.state;
        ps.pixel[slot] = pix;
        // No hit until the path is traced: a slot's previous path's hit must
        // never reach shading (it may index another scene's arrays).
        ps.hit[slot] = make_uint4(0x7f800000u, kMissId, kMissId, 0u);
        ps.active[pos] = slot | (job.tag << kSlotBits);
    }
}


// ------------------------------------------------------ queue construction
//
// Queues are built without contended atomics: a producer kernel records, per
// queue position i, which output queues the item goes to (a bit mask) and
// counts them per block; a one-block plan kernel scans the per-block counts into
// per-block output offsets; a scatter kernel walks the same positions in the
// same block order and writes each item at offset + rank. Every device-wide
// counter is then touched once per bounce instead of once per wave, and queue
// order no longer depends on timing.

// Lanes' bits of `flags` counted per queue over a block's loop, one count per wave.
template <int NQ>
struct WaveCounts {
    uint32_t c[NQ];
    __device__ __forceinline__ WaveCounts()
    {
#pragma unroll
        for (int q = 0; q < NQ; ++q) c[q] = 0;
    }
    __device__ __forceinline__ void add(uint32_t flags)
    {
#pragma unroll
        for (int q = 0; q < NQ; ++q) c[q] += uint32_t(__popcll(__ballot((flags >> q) & 1u)));
    }
};

// Writes the block's per-queue totals to block_count[blockIdx.x * 8 + q].
template <int BLOCK, int NQ>
__device__ __forceinline__ void storeBlockCounts(const WaveCounts<NQ>& wc, uint32_t* block_count)
{
    __shared__ uint32_t s[BLOCK / 64][NQ];
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) s[wave][q] = wc.c[q];
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(NQ)) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < uint32_t(BLOCK / 64); ++w) sum += s[w][threadIdx.x];
        block_count[blockIdx.x * 8 + threadIdx.x] = sum;
    }
}

// Exclusive scan of block_count over `grid` blocks for each of NQ queues into
// block_off. concat: queues share one output array, queue q after queues < q.
// Totals go to totals[q]. One block of 1024 threads: each thread sums a run of
// consecutive blocks, waves scan their threads' sums with shuffles, and the 16
// wave totals are combined through LDS.
// kPer: the most blocks a thread covers, ceil(grid / 1024), fixed at compile time
// so the loads all go out at once.
template <int NQ, uint32_t kPer = kMaxGrid / 1024>
__device__ __forceinline__ void scanBlockCounts(const uint32_t* block_count, uint32_t* block_off, uint32_t grid,
                                                bool concat, uint32_t* totals)
{
    constexpr uint32_t kWaves = 1024 / 64;
    __shared__ uint32_t wsum[kWaves][NQ];
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
    const uint32_t per = (grid + 1023) / 1024;
    const uint32_t b0 = min(grid, t * per), b1 = min(grid, b0 + per);
    // All of the thread's counts are loaded at once (unused entries read block 0
    // and count as zero), so the loads overlap instead of forming a chain.
    uint32_t cnt[NQ][kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const bool ok = b0 + k < b1;
        const uint32_t b = ok ? b0 + k : 0u;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const uint32_t x = block_count[b * 8 + q];
            cnt[q][k] = ok ? x : 0u;
        }
    }
    uint32_t local[NQ], incl[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) v += cnt[q][k];
        local[q] = v;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off);
            if (lane >= off) v += u;
        }
        incl[q] = v;
        if (lane == 63) wsum[wave][q] = v;
    }
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t before = 0, all = 0;
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t v = wsum[w][q];
            before += w < wave ? v : 0;
            all += v;
        }
        uint32_t run = (concat ? base : 0) + before + incl[q] - local[q];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            if (b0 + k < b1) block_off[(b0 + k) * 8 + q] = run;
            run += cnt[q][k];
        }
        if (t == 0) totals[q] = all;
        base += all;
    }
    __syncthreads();
}

// Rank of this lane among the block's lanes with bit q of `flags`, for one
// loop iteration; total[q] receives the iteration's count. All threads call it.
template <int BLOCK, int NQ>
__device__ __forceinline__ void blockRanks(uint32_t flags, uint32_t (&rank)[NQ], uint32_t (&total)[NQ])
{
    __shared__ uint32_t s[BLOCK / 64][NQ];
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const unsigned long long m = __ballot((flags >> q) & 1u);
        rank[q] = uint32_t(__popcll(m & below));
        if (lane == 0) s[wave][q] = uint32_t(__popcll(m));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t before = 0, all = 0;
        for (uint32_t w = 0; w < uint32_t(BLOCK / 64); ++w) {
            const uint32_t v = s[w][q];
            before += w < wave ? v : 0;
            all += v;
        }
        rank[q] += before;
        total[q] = all;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- tag sort
//
// Split tracing (a queue of several scenes): the extension queue's positions
// grouped by the scene tag of their entries, a counting sort over kTagGrid
// blocks that each own a contiguous range of the queue. Untagged positions come
// first, for the one-ray-per-lane kernels; each scene's rays then form one run
// for kTraceGrouped, however the scenes' paths lie in the queue. Within a tag,
// a block's positions keep their order in runs of at most a wave, so
// neighbouring lanes still read neighbouring slots.

// The queue range [lo, hi) of tag-sort block b.
__device__ __forceinline__ void tagRange(uint32_t n, uint32_t b, uint32_t& lo, uint32_t& hi)
{
    const uint32_t per = ((n + kTagGrid - 1) / kTagGrid + kTagBlock - 1) / kTagBlock * kTagBlock;
    lo = min(n, b * per);
    hi = min(n, lo + per);
}

// For each distinct tag among the wave's valid lanes, in turn: calls
// f(tag, lanes with it) on every lane (the mask is wave-uniform).
template <class F>
__device__ __forceinline__ void forEachTag(bool valid, uint32_t tag, F&& f)
{
    unsigned long long rest = __ballot(valid);
    while (rest) {
        const int first = __ffsll(static_cast<long long>(rest)) - 1;
        const uint32_t t = __shfl(tag, first);
        const unsigned long long m = __ballot(valid && tag == t);
        f(t, m);
        rest &= ~m;
    }
}

__global__ __launch_bounds__(kTagBlock) void kTagCount(PathState ps)
{
    __shared__ uint32_t hist[kMaxTags];
    const uint32_t n = ps.ctr->n_active + pendingGeneration(ps.ctr).n_extra;
    uint32_t lo, hi;
    tagRange(n, blockIdx.x, lo, hi);
    if (threadIdx.x < kMaxTags) hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t lane = __lane_id();
    for (uint32_t base = lo; base < hi; base += kTagBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint32_t tag = valid ? ps.active[i] >> kSlotBits : 0u;
        forEachTag(valid, tag, [&](uint32_t t, unsigned long long m) {
            if (lane == uint32_t(__ffsll(static_cast<long long>(m)) - 1)) atomicAdd(&hist[t], uint32_t(__popcll(m)));
        });
    }
    __syncthreads();
    if (threadIdx.x < kMaxTags) ps.tag_count[threadIdx.x * kTagGrid + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of tag_count in tag-major order (tag t's blocks after every
// block of the tags below it), in place; the untagged total to the counters.
__global__ __launch_bounds__(1024) void kTagPlan(PathState ps)
{
    constexpr uint32_t kAll = kMaxTags * kTagGrid, kPer = kAll / 1024, kWaves = 1024 / 64;
    static_assert(kAll % 1024 == 0 && kTagGrid % kPer == 0, "each thread scans whole runs of one tag");
    __shared__ uint32_t wsum[kWaves];
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
    uint32_t v[kPer];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        v[k] = ps.tag_count[t * kPer + k];
        local += v[k];
    }
    uint32_t incl = local;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - local;
    for (uint32_t w = 0; w < wave; ++w) run += wsum[w];
    // The thread whose entries start tag 1 begins at the untagged total.
    if (t * kPer == kTagGrid) ps.ctr->n_untagged = run;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        ps.tag_count[t * kPer + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kTagBlock) void kTagScatter(PathState ps)
{
    __shared__ uint32_t run[kMaxTags];
    const uint32_t n = ps.ctr->n_active + pendingGeneration(ps.ctr).n_extra;
    uint32_t lo, hi;
    tagRange(n, blockIdx.x, lo, hi);
    if (threadIdx.x < kMaxTags) run[threadIdx.x] = ps.tag_count[threadIdx.x * kTagGrid + blockIdx.x];
    __syncthreads();
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = lo; base < hi; base += kTagBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint32_t tag = valid ? ps.active[i] >> kSlotBits : 0u;
        forEachTag(valid, tag, [&](uint32_t t, unsigned long long m) {
            const int first = __ffsll(static_cast<long long>(m)) - 1;
            uint32_t at = 0;
            if (lane == uint32_t(first)) at = atomicAdd(&run[t], uint32_t(__popcll(m)));
            at = __shfl(at, first);
            if (valid && tag == t) ps.bytag[at + uint32_t(__popcll(m & below))] = i;
        });
    }
}

// ------------------------------------------------------------------ extend

__device__ __forceinline__ uint32_t materialBin(const DevScene& sc, uint32_t obj)
{
    if (obj == kMissId) return 0;
    return sc.materials[sc.objects[obj].material].kind + 1;
}

// SPLIT: only the untagged positions, which the tag sort put first in `bytag`;
// rays whose queue entries carry a scene tag are left to kTraceGrouped (their
// keys too; the block counts then come from kCountBins).
template <int BLOCK, int STACK, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void kExtend(PathState ps)
{
    __shared__ int32_t stack[STACK * BLOCK];
    // The queue holds the surviving (and recycled) paths and, after them, this
    // bounce's new paths in free slots.
    const uint32_t n = SPLIT ? ps.ctr->n_untagged : ps.ctr->n_active + pendingGeneration(ps.ctr).n_extra;
    WaveCounts<kBins> wc;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t i = SPLIT && k < n ? ps.bytag[k] : k;
        uint32_t flags = 0;
        const uint32_t entry = k < n ? ps.active[i] : 0u;
        const uint32_t slot = entry & kSlotMask;
        if (k < n) {
            const uint32_t j = __float_as_uint(ps.prev[slot].w);
            const float4 o = ps.org[slot], d = ps.dir[slot];
            const DevScene& sc = ps.jobs[j].scene;
            const RayF r{o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, __builtin_huge_valf()};
            HitF h;
            traceRay<BLOCK, false>(sc, r, h, LdsStack<BLOCK>{stack + threadIdx.x});
            ps.hit[slot] = make_uint4(__float_as_uint(h.t), h.object, h.prim, __float_as_uint(h.b1));
            ps.hit_b2[slot] = h.b2;
            flags = 1u << materialBin(sc, h.object);
            ps.key[i] = flags;
        }
        wc.add(flags);
    }
    storeBlockCounts<BLOCK, kBins>(wc, ps.block_count);
}

template <uint32_t PER>
__global__ __launch_bounds__(1024) void kPlanBins(PathState ps, uint32_t grid)
{
    __shared__ uint32_t totals[kBins];
    // The counter block, read once up front (no kernel writes it meanwhile), so
    // its load overlaps the scan instead of a chain of read-modify-writes after it.
    Counters c{};
    if (threadIdx.x == 0) c = *ps.ctr;
    scanBlockCounts<kBins, PER>(ps.block_count, ps.block_off, grid, true, totals);
    if (threadIdx.x == 0) {
        // The bookkeeping of the last finish and of this bounce's generation, both
        // already used by kGenerate and kExtend (pendingGeneration) and applied
        // here, where one thread runs after them anyway.
        const Generation g = pendingGeneration(&c);
        c.paths_done += c.n_term;
        c.shadow_rays += c.n_shadow;
        c.n_term = 0;
        c.n_shadow = 0;
        c.n_recycle = 0;
        c.n_free = g.free_top - g.n_extra;
        c.n_active += g.n_extra;
        // Appending pops from a sorted stack (sorted this bounce if it was not),
        // and what is left stays sorted.
        if (g.n_extra > 0) c.free_sorted = 1;
        c.work_next += g.n_new;
        c.camera_rays += g.n_new;
        c.extend_rays += c.n_active;
        c.n_in = c.n_active;
        for (int b = 0; b < kBins; ++b) c.bin_count[b] = totals[b];
        *ps.ctr = c;
    }
}

// Counting sort of the extended paths by material, so waves of the shade kernel
// mostly run one material's code. Only the shade kernel reads in this order:
// `binned` holds positions in `active`, and the queues the bounce produces keep
// the order of `active`, so samples of one pixel stay near each other from
// bounce to bounce.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kScatterBins(PathState ps)
{
    const uint32_t n = ps.ctr->n_in;
    uint32_t run[kBins];
#pragma unroll
    for (int b = 0; b < kBins; ++b) run[b] = ps.block_off[blockIdx.x * 8 + b];
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t flags = i < n ? ps.key[i] : 0u;
        uint32_t rank[kBins], total[kBins];
        blockRanks<BLOCK, kBins>(flags, rank, total);
        if (i < n) {
            uint32_t pos = 0;
#pragma unroll
            for (int b = 0; b < kBins; ++b)
                if ((flags >> b) & 1u) pos = run[b] + rank[b];
            ps.binned[pos] = i;
        }
#pragma unroll
        for (int b = 0; b < kBins; ++b) run[b] += total[b];
    }
}

// ------------------------------------------------------------------- shade

struct SurfaceHit {
    F3 p, n, ns; // point, geometric and shading normals facing the ray
    bool front;
    float u, v;
};

__device__ __forceinline__ SurfaceHit surfaceAt(const DevScene& sc, const uint4& h, float b2, F3 ro, F3 rd,
                                                bool want_uv)
{
    SurfaceHit s;
    const uint32_t obj = h.y;
    const DevObject& o = sc.objects[obj];
    const float t = __uint_as_float(h.x);
    F3 n, ns;
    if (o.kind == kObjSphere) {
        // reference geometry/sphere.cpp:119-134 (point, normal) and :137-147 (uv)
        const double* sp = sc.spheres + 4 * size_t(obj);
        const double px = double(ro.x) + double(rd.x) * double(t);
        const double py = double(ro.y) + double(rd.y) * double(t);
        const double pz = double(ro.z) + double(rd.z) * double(t);
        const double ex = px - sp[0], ey = py - sp[1], ez = pz - sp[2];
        const double il = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
        n = f3(float(ex * il), float(ey * il), float(ez * il));
        ns = n;
        s.p = f3(float(px), float(py), float(pz));
        s.u = s.v = 0.0f;
        if (want_uv) {
            s.u = (atan2f(-n.z, n.x) + kPiF) / (2.0f * kPiF);
            s.v = acosf(fminf(fmaxf(-n.y, -1.0f), 1.0f)) / kPiF;
        }
    } else {
        // reference mesh_geometry.cpp:77-98, mesh_instance.cpp:32-60
        const DevTriShading& tr = sc.shading[o.tri_shading_base + h.z];
        const float b1 = __uint_as_float(h.w), b0 = 1.0f - b1 - b2;
        const F3 p0 = f3(tr.p0[0], tr.p0[1], tr.p0[2]), p1 = f3(tr.p1[0], tr.p1[1], tr.p1[2]),
                 p2 = f3(tr.p2[0], tr.p2[1], tr.p2[2]);
        const F3 pl = p0 + (p1 - p0) * b1 + (p2 - p0) * b2;
        const F3 fn = f3(tr.fn[0], tr.fn[1], tr.fn[2]);
        F3 snl = fn;
        if (o.smooth) {
            const F3 bl = f3(tr.n0[0], tr.n0[1], tr.n0[2]) * b0 + f3(tr.n1[0], tr.n1[1], tr.n1[2]) * b1 +
                          f3(tr.n2[0], tr.n2[1], tr.n2[2]) * b2;
            if (len2(bl) >= 1e-12f) snl = normalize(bl);
        }
        s.u = b0 * tr.uv[0] + b1 * tr.uv[2] + b2 * tr.uv[4];
        s.v = b0 * tr.uv[1] + b1 * tr.uv[3] + b2 * tr.uv[5];
        const float* W = o.to_world;
        const float* I = o.to_object;
        s.p = f3(W[0] * pl.x + W[1] * pl.y + W[2] * pl.z + W[3], W[4] * pl.x + W[5] * pl.y + W[6] * pl.z + W[7],
                 W[8] * pl.x + W[9] * pl.y + W[10] * pl.z + W[11]);
        n = normalize(f3(I[0] * fn.x + I[4] * fn.y + I[8] * fn.z, I[1] * fn.x + I[5] * fn.y + I[9] * fn.z,
                         I[2] * fn.x + I[6] * fn.y + I[10] * fn.z));
        if (o.smooth) {
            const F3 sw = f3(I[0] * snl.x + I[4] * snl.y + I[8] * snl.z, I[1] * snl.x + I[5] * snl.y + I[9] * snl.z,
                             I[2] * snl.x + I[6] * snl.y + I[10] * snl.z);
            ns = isNormalF(len2(sw)) ? normalize(sw) : n;
        } else {
            ns = n;
        }
    }
    s.front = dot(rd, n) < 0.0f;
    s.n = s.front ? n : -n;
    s.ns = s.front ? ns : -ns;
    return s;
}

// Cone of an area light's sphere seen from p (reference light/area_light.cpp:31-51).
struct ConeF {
    F3 axis;
    float dist, tangent2, cos_max, omc;
};
__device__ __forceinline__ bool coneTowards(const DevLight& l, F3 p, ConeF& c)
{
    const F3 tc = f3(float(l.position[0] - double(p.x)), float(l.position[1] - double(p.y)),
                     float(l.position[2] - double(p.z)));
    const float d2 = len2(tc);
    const float r = float(l.radius);
    const float r2 = r * r;
    if (d2 <= r2) return false;
    const float d = sqrtf(d2);
    const float s2 = r2 / d2;
    const float cm = sqrtf(1.0f - s2);
    const float omc = s2 / (1.0f + cm);
    if (!isNormalF(omc)) return false;
    c.axis = tc / d;
    c.dist = d;
    c.tangent2 = d2 - r2;
    c.cos_max = cm;
    c.omc = omc;
    return true;
}

// Density of choosing direction `dir` from p by sampling light l (reference
// light/area_light.cpp:84-94, sky_light.hpp, point_light.hpp).
__device__ __forceinline__ float lightPdf(const DevLight& l, F3 p, F3 dir)
{
    if (l.kind == 0u) return 0.0f;
    if (l.kind == 2u) return kUniformSpherePdf;
    ConeF c;
    if (!coneTowards(l, p, c)) return 0.0f;
    if (dot(dir, c.axis) < c.cos_max) return 0.0f;
    return 1.0f / (2.0f * kPiF * c.omc);
}

__device__ __forceinline__ F3 skyRadiance(const DevLight& l, F3 dir)
{
    const float h = 0.5f * (dir.y + 1.0f);
    return f3(l.a[0], l.a[1], l.a[2]) * (1.0f - h) + f3(l.b[0], l.b[1], l.b[2]) * h;
}

struct LightSampleF {
    F3 dir, radiance;
    float dist, pdf;
    bool delta;
};

// reference light/point_light.cpp:7-21, area_light.cpp:53-82, sky_light.cpp:17-24
__device__ __forceinline__ bool sampleLight(const DevLight& l, F3 p, float ux, float uy, LightSampleF& s)
{
    if (l.kind == 0u) {
        const F3 to = f3(float(l.position[0] - double(p.x)), float(l.position[1] - double(p.y)),
                         float(l.position[2] - double(p.z)));
        const float d2 = len2(to);
        if (d2 == 0.0f) return false;
        const float d = sqrtf(d2);
        s.dir = to / d;
        s.radiance = f3(l.a[0], l.a[1], l.a[2]) / d2;
        s.dist = d;
        s.pdf = 1.0f;
        s.delta = true;
        return true;
    }
    if (l.kind == 2u) {
        s.dir = uniformSphere(ux, uy);
        s.radiance = skyRadiance(l, s.dir);
        s.dist = __builtin_huge_valf();
        s.pdf = kUniformSpherePdf;
        s.delta = false;
        return true;
    }
    ConeF c;
    if (!coneTowards(l, p, c)) return false;
    const float omc = ux * c.omc;
    const float ct = 1.0f - omc;
    const float s2 = omc * (2.0f - omc);
    const float st = sqrtf(s2);
    float sn, cs;
    sincosf(2.0f * kPiF * uy, &sn, &cs);
    const Frame f = frameFrom(c.axis);
    s.dir = toWorld(f, f3(st * cs, st * sn, ct));
    const float r = float(l.radius);
    const float root = sqrtf(fmaxf(0.0f, r * r - c.dist * c.dist * s2));
    s.dist = c.tangent2 / (c.dist * ct + root);
    s.radiance = f3(l.a[0], l.a[1], l.a[2]);
    s.pdf = 1.0f / (2.0f * kPiF * c.omc);
    s.delta = false;
    return true;
}

// One bounce of PathEstimator::radiance (reference path_estimator.cpp:153-185)
// for every extended path: emission with MIS, next-event estimation (the shadow
// ray is deferred to kShadow), BSDF sampling and Russian roulette.
constexpr uint32_t kQueueActive = 1u, kQueueShadow = 2u, kQueueTerm = 4u;

// Shades the path in `slot` and returns the queues it joins.
__device__ __forceinline__ uint32_t shadePath(const PathState& ps, uint32_t slot)
{
        const float4 o4 = ps.org[slot], d4 = ps.dir[slot], t4 = ps.thr[slot], p4 = ps.prev[slot];
        const uint32_t j = __float_as_uint(p4.w);
        const DevJob& job = ps.jobs[j];
        const DevScene& sc = job.scene;
        const F3 ro = f3(o4.x, o4.y, o4.z), rd = f3(d4.x, d4.y, d4.z), arr_o = f3(p4.x, p4.y, p4.z);
        const float arr_pdf = o4.w;
        const bool arr_spec = d4.w != 0.0f;
        F3 thr = f3(t4.x, t4.y, t4.z);
        // The path's radiance, read and written only by the paths whose radiance
        // changes here (most bounces add nothing); additions keep their order.
        F3 L = f3(0.0f, 0.0f, 0.0f);
        bool l_set = false;
        auto addL = [&](F3 x) {
            if (!l_set) {
                const float4 r4 = ps.rad[slot];
                L = f3(r4.x, r4.y, r4.z);
                l_set = true;
            }
            L = L + x;
        };
        const int depth = __float_as_int(t4.w);
        const uint4 h = ps.hit[slot];
        // Ids out of the scene's range (a hit that was never written for this
        // path) count as a miss rather than index past the scene's arrays.
        bool miss = h.y >= sc.num_objects;
        if (!miss) {
            const DevObject& ho = sc.objects[h.y];
            miss = ho.kind != kObjSphere && h.z >= ho.tri_count;
        }
        const float nl = float(sc.num_lights);
        // Estimators (reference estimator/radiance_estimator.hpp:36-38): the path
        // estimator weighs emission by MIS and adds next-event estimation and
        // roulette; the scatter-only one (scatter_only_estimator.cpp:10-48) only
        // gathers the emission its BSDF samples run into; the normals view
        // (normals_estimator.cpp:15-40) colours the first hit by its shading normal.
        const uint32_t est = job.estimator;
        const bool mis = est == AB_ESTIMATOR_PATH;
        bool alive = false, shadow = false;
        if (job.cancelled) {
            // terminated without gathering
        } else if (miss) {
            if (est == AB_ESTIMATOR_NORMALS) {
                // White straight down to (0.5, 0.7, 1) straight up, on the normalised direction.
                const float up = 0.5f * (normalize(rd).y + 1.0f);
                L = f3(1.0f, 1.0f, 1.0f) * (1.0f - up) + f3(0.5f, 0.7f, 1.0f) * up;
                l_set = true;
            } else if (sc.sky_light >= 0) {
                const DevLight& sky = sc.lights[sc.sky_light];
                F3 e = skyRadiance(sky, rd);
                if (mis && !isBlack(e) && !arr_spec) e = e * powerHeuristic(arr_pdf, (1.0f / nl) * kUniformSpherePdf);
                addL(thr * e);
            }
        } else {
            const DevObject& obj = sc.objects[h.y];
            const DevMaterial* m = &sc.materials[obj.material];
            const bool uv_needed = est != AB_ESTIMATOR_NORMALS && (m->kind == 0u || m->kind == 3u);
            const SurfaceHit s = surfaceAt(sc, h, ps.hit_b2[slot], ro, rd, uv_needed);
            if (est == AB_ESTIMATOR_NORMALS) {
                L = (s.ns + f3(1.0f, 1.0f, 1.0f)) * 0.5f;
                l_set = true;
            } else if (obj.light >= 0 && s.front) {
                const DevLight& l = sc.lights[obj.light];
                if (l.kind == 1u) {
                    F3 e = f3(l.a[0], l.a[1], l.a[2]);
                    if (mis && !isBlack(e) && !arr_spec)
                        e = e * powerHeuristic(arr_pdf, (1.0f / nl) * lightPdf(l, arr_o, rd));
                    addL(thr * e);
                }
            }
            if (est != AB_ESTIMATOR_NORMALS && depth < int(job.max_depth)) {
                Rng rng;
                rng.state = ps.rng[slot];
                rng.inc = (mixBits(ps.pixel[slot]) << 1u) | 1u;
                if (m->kind == 3u) {
                    const float w = patternWeight(s.p, m->mix_scale, m->mix_veins);
                    m = &sc.materials[w < m->mix_threshold ? m->first : m->second];
                }
                Scatter sct;
                sct.frame = frameFrom(s.ns);
                sct.geo = s.n;
                const float a = m->roughness * m->roughness;
                if (m->kind == 1u) {
                    sct.lobe = Lobe{kConductor, f3(m->a[0], m->a[1], m->a[2]), f3(m->b[0], m->b[1], m->b[2]), a, 1.0f};
                } else if (m->kind == 2u) {
                    sct.lobe = Lobe{kDielectric, f3(0, 0, 0), f3(0, 0, 0), a, s.front ? m->ior : 1.0f / m->ior};
                } else {
                    sct.lobe = Lobe{kLambert, textureAt(*m, s.p, s.u, s.v), f3(0, 0, 0), 0.0f, 1.0f};
                }
                const F3 wo = -rd;
                // Next-event estimation (path_estimator.cpp:73-102): three draws always.
                const float ul = mis ? rng.uniform() : 0.0f, lx = mis ? rng.uniform() : 0.0f,
                            ly = mis ? rng.uniform() : 0.0f;
                if (mis && sc.num_lights > 0) {
                    const uint32_t li = min(uint32_t(ul * nl), sc.num_lights - 1);
                    const DevLight& l = sc.lights[li];
                    LightSampleF ls;
                    if (sampleLight(l, s.p, lx, ly, ls) && !isBlack(ls.radiance)) {
                        const F3 f = sct.eval(wo, ls.dir);
                        if (!isBlack(f)) {
                            // Stop short of the light by a relative 1e-3 (at least 1e-3):
                            // more than the FP32 ray-origin offset moves the distance.
                            const float tmax = isinf(ls.dist) ? ls.dist : ls.dist - fmaxf(ls.dist * 1e-3f, 1e-3f);
                            const float lpdf = (1.0f / nl) * ls.pdf;
                            const float w = ls.delta ? 1.0f : powerHeuristic(lpdf, sct.pdf(wo, ls.dir));
                            const float ct = fabsf(dot(ls.dir, s.ns));
                            const F3 contrib = thr * f * ls.radiance * (ct * w / lpdf);
                            if (tmax > 0.0f) {
                                const F3 so = offsetRay(s.p, dot(ls.dir, s.n) >= 0.0f ? s.n : -s.n);
                                ps.sh_org[slot] = make_float4(so.x, so.y, so.z, tmax);
                                ps.sh_dir[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, 0.0f);
                                ps.sh_c[slot] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
                                shadow = true;
                            } else {
                                addL(contrib);
                            }
                        }
                    }
                }
                const float uc = rng.uniform(), ux = rng.uniform(), uy = rng.uniform();
                Sample smp;
                if (sct.sample(wo, uc, ux, uy, smp)) {
                    thr = thr * smp.weight;
                    if (!isBlack(thr)) {
                        alive = true;
                        if (mis && depth >= int(job.roulette_start)) {
                            const float surv = fminf(0.95f, maxc(thr));
                            if (rng.uniform() >= surv) alive = false;
                            else thr = thr / surv;
                        }
                    }
                }
                ps.rng[slot] = rng.state;
                if (alive) {
                    const F3 no = offsetRay(s.p, dot(smp.dir, s.n) >= 0.0f ? s.n : -s.n);
                    ps.org[slot] = make_float4(no.x, no.y, no.z, smp.pdf);
                    ps.dir[slot] = make_float4(smp.dir.x, smp.dir.y, smp.dir.z, smp.specular ? 1.0f : 0.0f);
                    ps.thr[slot] = make_float4(thr.x, thr.y, thr.z, __int_as_float(depth + 1));
                    ps.prev[slot] = make_float4(s.p.x, s.p.y, s.p.z, p4.w);
                }
            }
        }
        if (l_set) ps.rad[slot] = make_float4(L.x, L.y, L.z, 0.0f);
        return (alive ? kQueueActive : kQueueTerm) | (shadow ? kQueueShadow : 0u);
}

// Shades in material order; the queue flags go to the path's position in
// `active`.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kShade(PathState ps)
{
    const uint32_t n = ps.ctr->n_in;
    for (uint32_t k = gridStart(); k < n; k += gridStride()) {
        const uint32_t i = ps.binned[k];
        ps.key[i] = shadePath(ps, ps.active[i] & kSlotMask);
    }
}

// Counts the shaded paths per output queue and block, in `active` order (TAGGED:
// in the tag sort's order, so the queues built from it come out sorted by tag).
template <int BLOCK, bool TAGGED>
__global__ __launch_bounds__(BLOCK) void kCountShade(PathState ps)
{
    const uint32_t n = ps.ctr->n_in;
    uint32_t lo, hi;
    blockSpan<BLOCK>(n, lo, hi);
    WaveCounts<3> wc;
    for (uint32_t base = lo; base < hi; base += BLOCK) {
        const uint32_t k = base + threadIdx.x;
        wc.add(k < hi ? ps.key[TAGGED ? ps.bytag[k] : k] : 0u);
    }
    storeBlockCounts<BLOCK, 3>(wc, ps.block_count);
}

// Sums of block_count[b * 8 + q] over the blocks b < limit, for each q < NQ, to
// every thread of the block. Each block of a scatter kernel sums the counts of
// the blocks before it itself (at most kMaxGrid x NQ words, from L2), so no
// one-block scan kernel has to run between the count and the scatter.
template <int BLOCK, int NQ>
__device__ __forceinline__ void countsBefore(const uint32_t* block_count, uint32_t limit, uint32_t (&sum)[NQ])
{
    __shared__ uint32_t s[BLOCK / 64][NQ];
    uint32_t v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = 0;
    for (uint32_t b = threadIdx.x; b < limit; b += BLOCK) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] += block_count[b * 8 + q];
    }
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off);
        if (__lane_id() == 0) s[wave][q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < uint32_t(BLOCK / 64); ++w) t += s[w][q];
        sum[q] = t;
    }
    __syncthreads();
}

// Writes the shaded paths into the next extension queue and the shadow and
// termination queues, keeping their order in `active` (TAGGED: in the tag
// sort's order; the shadow queue then holds its untagged entries first): block
// b writes its range's items after those of the blocks before it. When
// the work left covers every path of the bounce, terminated paths keep their
// positions too (recycling: the next bounce's first new samples take them
// over), so the next extension queue is this one.
template <int BLOCK, bool TAGGED>
__global__ __launch_bounds__(BLOCK) void kScatterShade(PathState ps)
{
    const Counters* cc = ps.ctr;
    const uint32_t n = cc->n_in;
    const bool recycle = cc->work_total - cc->work_next >= n;
    uint32_t run[3];
    countsBefore<BLOCK, 3>(ps.block_count, blockIdx.x, run);
    if (blockIdx.x == 0) {
        // The queue sizes for the kernels after this one (no block of this one
        // reads them).
        uint32_t all[3];
        countsBefore<BLOCK, 3>(ps.block_count, gridDim.x, all);
        if (threadIdx.x == 0) {
            Counters* c = ps.ctr;
            c->n_active = recycle ? n : all[0];
            c->n_shadow = all[1];
            c->n_term = all[2];
            c->n_recycle = recycle ? all[2] : 0u;
            if (TAGGED && cc->n_untagged >= n) c->sh_untagged = all[1];
        }
    }
    const uint32_t untagged = TAGGED ? cc->n_untagged : 0u;
    uint32_t lo, hi;
    blockSpan<BLOCK>(n, lo, hi);
    for (uint32_t base = lo; base < hi; base += BLOCK) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t i = TAGGED && k < hi ? ps.bytag[k] : k;
        const uint32_t flags = k < hi ? ps.key[i] : 0u;
        uint32_t rank[3], total[3];
        blockRanks<BLOCK, 3>(flags, rank, total);
        // The shadow entries of the untagged positions, which come first.
        if (TAGGED && k == untagged && k < hi) ps.ctr->sh_untagged = run[1] + rank[1];
        if (k < hi) {
            const uint32_t entry = ps.active[i]; // the slot and its scene tag
            if (recycle) ps.active_next[i] = entry;
            else if (flags & kQueueActive) ps.active_next[run[0] + rank[0]] = entry;
            if (flags & kQueueShadow) ps.shadowq[run[1] + rank[1]] = entry;
            if (flags & kQueueTerm) ps.termq[run[2] + rank[2]] = i;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) run[q] += total[q];
    }
}

// ---------------------------------------------------------- free-slot sort

// The free stack rebuilt so that it pops slots in ascending order, before a
// generation that appends paths to the queue: the appended paths then take
// ascending slots, so neighbouring lanes read neighbouring slots. A finish
// pushes its freed slots in queue order on top of older free slots, so
// without this, paths appended after the queue ran short of work (a render's
// last samples, a server between arrivals) would take slots spread over the
// whole state arrays, and recycling would keep them there. Runs only when the
// stack is unsorted and paths are appended; the slots' free flags give the set.
__device__ __forceinline__ bool freeSortNeeded(const Counters* c)
{
    return !c->free_sorted && pendingGeneration(c).n_extra > 0;
}

// The slot range [lo, hi) of free-sort block b.
__device__ __forceinline__ void freeRange(uint32_t cap, uint32_t b, uint32_t& lo, uint32_t& hi)
{
    const uint32_t per = ((cap + kFreeGrid - 1) / kFreeGrid + 255) / 256 * 256;
    lo = min(cap, b * per);
    hi = min(cap, lo + per);
}

__global__ __launch_bounds__(256) void kFreeCount(PathState ps)
{
    if (!freeSortNeeded(ps.ctr)) return;
    uint32_t lo, hi;
    freeRange(ps.capacity, blockIdx.x, lo, hi);
    WaveCounts<1> wc;
    for (uint32_t base = lo; base < hi; base += 256) {
        const uint32_t s = base + threadIdx.x;
        wc.add(s < hi && ps.slot_free[s] ? 1u : 0u);
    }
    storeBlockCounts<256, 1>(wc, ps.block_count);
}

__global__ __launch_bounds__(256) void kFreeScatter(PathState ps)
{
    if (!freeSortNeeded(ps.ctr)) return;
    const uint32_t free_top = pendingGeneration(ps.ctr).free_top;
    uint32_t run[1];
    countsBefore<256, 1>(ps.block_count, blockIdx.x, run);
    uint32_t lo, hi;
    freeRange(ps.capacity, blockIdx.x, lo, hi);
    for (uint32_t base = lo; base < hi; base += 256) {
        const uint32_t s = base + threadIdx.x;
        const uint32_t flag = s < hi && ps.slot_free[s] ? 1u : 0u;
        uint32_t rank[1], total[1];
        blockRanks<256, 1>(flag, rank, total);
        // Stored from the top down, so the stack pops the lowest slot first.
        const uint32_t r = run[0] + rank[0];
        if (flag && r < free_top) ps.freeq[free_top - 1 - r] = s;
        run[0] += total[0];
    }
}

// ------------------------------------------------------------------ shadow

// Occlusion of the deferred light samples (reference Scene::occluded,
// scene/scene.cpp:157-165); unoccluded contributions join the path's radiance.
// SPLIT: only the untagged entries, which come first in the shadow queue; rays
// whose queue entries carry a scene tag are left to kTraceGrouped.
template <int BLOCK, int STACK, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void kShadow(PathState ps)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const uint32_t n = SPLIT ? ps.ctr->sh_untagged : ps.ctr->n_shadow;
    for (uint32_t i = gridStart(); i < n; i += gridStride()) {
        const uint32_t entry = ps.shadowq[i];
        const uint32_t slot = entry & kSlotMask;
        const uint32_t j = __float_as_uint(ps.prev[slot].w);
        const float4 o = ps.sh_org[slot], d = ps.sh_dir[slot];
        const RayF r{o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, o.w};
        HitF h;
        if (!traceRay<BLOCK, true>(ps.jobs[j].scene, r, h, LdsStack<BLOCK>{stack + threadIdx.x})) {
            const float4 cc = ps.sh_c[slot];
            float4 rr = ps.rad[slot];
            rr.x += cc.x;
            rr.y += cc.y;
            rr.z += cc.z;
            ps.rad[slot] = rr;
        }
    }
}

// ------------------------------------------------- extend and shadow, lane refill

// For queues whose rays traverse mesh instances: the extension and shadow rays
// traced as the ray-query kernel traces them. Each one-wave block owns a
// contiguous slice of the queue, and a lane whose ray has finished takes the
// next ray of the slice once kQueueRefill lanes are free, so a wave no longer
// waits for its longest ray. RayQuery visits the same nodes and primitives as
// traceRay, so the hits and shadow results are those of kExtend and kShadow. The
// material keys are counted afterwards by kCountBins, over the block partition
// kScatterBins uses.
constexpr uint32_t kQueueRefill = 16; // free lanes that trigger a refill
constexpr uint32_t kQueueSteps = 16;  // traversal steps between refill checks

// Up to kQueueSteps steps of the lanes holding a ray (called by those lanes
// only). The wave leaves early once kQueueRefill lanes are free or done while its
// slice has rays left to start (`more`, wave-uniform), so lanes whose rays end
// early (culled at the root, or beyond the sphere hit) take new ones instead of
// waiting out the chunk. The refill kernels' variants with the FP64 sphere test
// run whole chunks instead: at their 80 VGPRs this check makes the compiler spill.
template <class Query>
__device__ __forceinline__ void stepQueue(Query& q, const DevScene& sc, LdsStack<64> stk, bool more)
{
    for (uint32_t k = 0; k < kQueueSteps && !q.done; ++k) {
        q.step(sc, stk);
        if (more && uint32_t(__popcll(__ballot(!q.done))) <= 64 - kQueueRefill) break;
    }
}

// Writes the finished query of the ray in `slot` at queue position `pos`.
template <bool ANY, class Query>
__device__ __forceinline__ void finishQuery(const PathState& ps, const DevScene& sc, const Query& q, uint32_t slot,
                                            uint32_t pos)
{
    if (ANY) {
        if (q.best.object == kMissId) {
            const float4 cc = ps.sh_c[slot];
            float4 rr = ps.rad[slot];
            rr.x += cc.x;
            rr.y += cc.y;
            rr.z += cc.z;
            ps.rad[slot] = rr;
        }
    } else {
        const HitF h = q.result();
        ps.hit[slot] = make_uint4(__float_as_uint(h.t), h.object, h.prim, __float_as_uint(h.b1));
        ps.hit_b2[slot] = h.b2;
        ps.key[pos] = 1u << materialBin(sc, h.object);
    }
}

template <bool ANY>
__device__ __forceinline__ RayF queuedRay(const PathState& ps, uint32_t slot)
{
    const float4 o = ANY ? ps.sh_org[slot] : ps.org[slot];
    const float4 d = ANY ? ps.sh_dir[slot] : ps.dir[slot];
    return RayF{o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, ANY ? o.w : __builtin_huge_valf()};
}

// First pass over a kind-split scene (`sph` holds its sphere-only TLAS), one ray
// per lane over the queue the refill kernel then traces: a closest-hit ray
// leaves its nearest sphere hit (or a miss) in the path's hit, which the
// instance pass starts from; a shadow ray a sphere occludes gets an empty range
// and no contribution, so the instance pass adds nothing for it. The objects,
// materials and arrays are the scene's own, so ids and keys are unchanged.
template <int BLOCK, int STACK, bool ANY>
__global__ __launch_bounds__(BLOCK) void kSpherePass(PathState ps, DevScene sph)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const Counters* c = ps.ctr;
    const uint32_t n = ANY ? c->n_shadow : c->n_active + pendingGeneration(c).n_extra;
    const LdsStack<BLOCK> stk{stack + threadIdx.x};
    for (uint32_t k = gridStart(); k < n; k += gridStride()) {
        const uint32_t slot = (ANY ? ps.shadowq[k] : ps.active[k]) & kSlotMask;
        const RayF r = queuedRay<ANY>(ps, slot);
        HitF h;
        const bool hit = traceRay<BLOCK, ANY>(sph, r, h, stk);
        if (ANY) {
            if (hit) {
                const float4 o = ps.sh_org[slot];
                ps.sh_org[slot] = make_float4(o.x, o.y, o.z, 0.0f);
                ps.sh_c[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        } else {
            ps.hit[slot] = make_uint4(__float_as_uint(h.t), h.object, h.prim, 0u);
        }
    }
}

// The same first pass over the tagged range of a tag-sorted queue (split
// tracing): each ray's scene comes from its entry's tag; a ray of a kind-split
// scene is traced over that scene's sphere TLAS, the rays of other scenes get a
// miss (a closest-hit ray must not start from its previous bounce's hit).
template <int BLOCK, int STACK, bool ANY>
__global__ __launch_bounds__(BLOCK) void kSpherePassTagged(PathState ps)
{
    __shared__ int32_t stack[STACK * BLOCK];
    const Counters* c = ps.ctr;
    const uint32_t first = ANY ? c->sh_untagged : c->n_untagged;
    const uint32_t n = ANY ? c->n_shadow : c->n_active + pendingGeneration(c).n_extra;
    const LdsStack<BLOCK> stk{stack + threadIdx.x};
    for (uint32_t k = first + gridStart(); k < n; k += gridStride()) {
        const uint32_t entry = ANY ? ps.shadowq[k] : ps.active[ps.bytag[k]];
        const uint32_t slot = entry & kSlotMask;
        DevScene sc = ps.tag_scenes[entry >> kSlotBits];
        HitF h{__builtin_huge_valf(), kMissId, kMissId, 0.0f, 0.0f};
        bool hit = false;
        if (sc.sphere_root != kNoChild) {
            sc.tlas_root = sc.sphere_root;
            hit = traceRay<BLOCK, ANY>(sc, queuedRay<ANY>(ps, slot), h, stk);
        }
        if (ANY) {
            if (hit) {
                const float4 o = ps.sh_org[slot];
                ps.sh_org[slot] = make_float4(o.x, o.y, o.z, 0.0f);
                ps.sh_c[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        } else {
            ps.hit[slot] = make_uint4(__float_as_uint(h.t), h.object, h.prim, 0u);
        }
    }
}

// Waves per SIMD the refill kernels ask the compiler for: with the FP64 sphere
// test 6 (80 VGPRs, which with the 24-entry stack class's 6 KB of LDS per wave
// leaves room for 6 waves), without it 7 (72 VGPRs).
template <bool SPHERES>
constexpr int refillWaves() { return SPHERES ? 6 : 7; }

// Every ray of the queue over one scene, passed as a kernel argument.
// SEEDED: the second pass of a kind-split scene (`sc` holds the instance-only
// TLAS): closest-hit queries start from the sphere hit kSpherePass left in the
// path's hit; shadow rays that pass found occluded arrive with an empty range.
template <int STACK, bool ANY, bool SPHERES, bool LEAF_IN_REG, bool SEEDED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(refillWaves<SPHERES>(), 8))) void kTraceQueue(
    PathState ps, DevScene sc)
{
    __shared__ int32_t stack[STACK * 64];
    const Counters* c = ps.ctr;
    const uint32_t n = ANY ? c->n_shadow : c->n_active + pendingGeneration(c).n_extra;
    const uint32_t per_lane = (n + 64 * gridDim.x - 1) / (64 * gridDim.x);
    const uint32_t slice = per_lane * 64;
    uint32_t next = xcdRemap(blockIdx.x, gridDim.x) * slice;
    const uint32_t end = next + slice < n ? next + slice : n;
    const uint32_t lane = __lane_id();
    const LdsStack<64> stk{stack + threadIdx.x};
    RayQuery<64, ANY, SPHERES, LEAF_IN_REG> q;
    constexpr uint32_t kNone = ~0u;
    uint32_t id = kNone, slot = 0;
    while (true) {
        const bool idle = id == kNone;
        const unsigned long long free_lanes = __ballot(idle);
        if (next < end && (__popcll(free_lanes) >= kQueueRefill || free_lanes == ~0ull)) {
            const uint32_t mine = next + uint32_t(__popcll(free_lanes & ((1ull << lane) - 1ull)));
            next += uint32_t(__popcll(free_lanes));
            if (idle && mine < end) {
                slot = (ANY ? ps.shadowq[mine] : ps.active[mine]) & kSlotMask;
                q.begin(sc, queuedRay<ANY>(ps, slot));
                if (SEEDED && !ANY) {
                    const uint4 h = ps.hit[slot];
                    q.seed(HitF{__uint_as_float(h.x), h.y, h.z, 0.0f, 0.0f});
                }
                id = mine;
            }
        }
        // Done when no lane holds a ray and the slice is used up.
        if (__ballot(id != kNone) == 0 && next >= end) break;
        if (id != kNone) {
            if constexpr (SPHERES) {
                for (uint32_t k = 0; k < kQueueSteps && !q.done; ++k) q.step(sc, stk);
            } else {
                stepQueue(q, sc, stk, next < end);
            }
            if (q.done) {
                finishQuery<ANY>(ps, sc, q, slot, id);
                id = kNone;
            }
        }
    }
}

// A value the whole wave holds the same, moved to a scalar register.
__device__ __forceinline__ uint32_t waveUniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ __forceinline__ T* waveUniform(T* p)
{
    const auto bits = reinterpret_cast<unsigned long long>(p);
    const unsigned long long lo = waveUniform(uint32_t(bits)), hi = waveUniform(uint32_t(bits >> 32));
    return reinterpret_cast<T*>((hi << 32) | lo);
}

// The scene of tag `tag`, read once into scalar registers: traversal then finds
// its arrays there, as it finds a kernel argument's, instead of loading them
// from memory at every step. `tag` must be the same in every lane.
__device__ __forceinline__ DevScene sceneOfTag(const PathState& ps, uint32_t tag)
{
    const DevScene& s = ps.tag_scenes[tag];
    DevScene u;
    u.nodes = waveUniform(s.nodes);
    u.tris = waveUniform(s.tris);
    u.leafobj = waveUniform(s.leafobj);
    u.objects = waveUniform(s.objects);
    u.spheres = waveUniform(s.spheres);
    u.shading = waveUniform(s.shading);
    u.materials = waveUniform(s.materials);
    u.lights = waveUniform(s.lights);
    u.tlas_root = int32_t(waveUniform(uint32_t(s.tlas_root)));
    u.num_objects = waveUniform(s.num_objects);
    u.num_lights = waveUniform(s.num_lights);
    u.sky_light = int32_t(waveUniform(uint32_t(s.sky_light)));
    u.stack_depth = waveUniform(s.stack_depth);
    u.has_spheres = waveUniform(s.has_spheres);
    u.has_meshes = waveUniform(s.has_meshes);
    u.tri_count = waveUniform(s.tri_count);
    u.wide_root = int32_t(waveUniform(uint32_t(s.wide_root)));
    u.wide_stack = waveUniform(s.wide_stack);
    u.sphere_root = int32_t(waveUniform(uint32_t(s.sphere_root)));
    u.mesh_root = int32_t(waveUniform(uint32_t(s.mesh_root)));
    u.sphere_stack = waveUniform(s.sphere_stack);
    u.mesh_stack = waveUniform(s.mesh_stack);
    return u;
}

// In a queue of several scenes, the rays whose queue entries carry a scene tag
// (the others go to kExtend / kShadow), one scene per wave at a time. The tag
// sort orders the extension queue by tag (through `bytag`), and the shadow queue
// is written in that order, so each scene's rays form one run after the
// untagged ones. When every lane is idle the wave starts a run at its slice's
// next ray and holds that run's scene in wave-uniform registers; later refills
// take only the following rays of the same tag and stop at the first ray of
// another. Traversal then reads one scene per wave from scalar registers, as
// kTraceQueue does with its kernel argument, instead of a scene per lane.
// SEEDED: a kind-split scene's run walks its instance-only TLAS, starting from
// the sphere hit kSpherePassTagged left (closest hits; shadow rays a sphere
// occluded arrive with an empty range); other scenes' runs walk their whole
// TLAS, and the pass left them a miss.
template <int STACK, bool ANY, bool SPHERES, bool LEAF_IN_REG, bool SEEDED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(refillWaves<SPHERES>(), 8))) void kTraceGrouped(
    PathState ps)
{
    __shared__ int32_t stack[STACK * 64];
    const Counters* c = ps.ctr;
    // The tagged entries: [first, n) of the tag-sorted order.
    const uint32_t first = ANY ? c->sh_untagged : c->n_untagged;
    const uint32_t n = ANY ? c->n_shadow : c->n_active + pendingGeneration(c).n_extra;
    const uint32_t per_lane = (n - first + 64 * gridDim.x - 1) / (64 * gridDim.x);
    const uint32_t slice = per_lane * 64;
    uint32_t next = first + xcdRemap(blockIdx.x, gridDim.x) * slice;
    const uint32_t end = next + slice < n ? next + slice : n;
    const uint32_t lane = __lane_id();
    const LdsStack<64> stk{stack + threadIdx.x};
    RayQuery<64, ANY, SPHERES, LEAF_IN_REG> q;
    constexpr uint32_t kNone = ~0u;
    uint32_t id = kNone, slot = 0;
    DevScene sc{};        // the run's scene (wave-uniform)
    uint32_t run_tag = 0; // ... and its tag
    while (true) {
        const bool idle = id == kNone;
        const unsigned long long free_lanes = __ballot(idle);
        const uint32_t n_free = uint32_t(__popcll(free_lanes));
        if (next < end && (n_free >= kQueueRefill || free_lanes == ~0ull)) {
            // Lane i looks at window position next + i: its queue position, entry
            // and the entry's scene tag (0: another kernel's ray).
            const uint32_t cnt = min(n_free, end - next);
            const uint32_t w_pos = lane < cnt ? (ANY ? next + lane : ps.bytag[next + lane]) : 0u;
            const uint32_t w_entry = lane < cnt ? (ANY ? ps.shadowq[w_pos] : ps.active[w_pos]) : 0u;
            const uint32_t w_tag = w_entry >> kSlotBits;
            uint32_t start = 0;
            bool fill = true;
            if (free_lanes == ~0ull) {
                const unsigned long long owned = __ballot(w_tag != 0u);
                if (owned == 0ull) {
                    next += cnt; // none of ours in the window
                    fill = false;
                } else {
                    start = uint32_t(__ffsll(static_cast<long long>(owned))) - 1;
                    run_tag = waveUniform(uint32_t(__shfl(w_tag, int(start))));
                    sc = sceneOfTag(ps, run_tag);
                    if (SEEDED && sc.mesh_root != kNoChild) {
                        sc.tlas_root = sc.mesh_root;
                        sc.has_spheres = 0;
                    }
                }
            }
            if (fill) {
                // The run goes on up to the first window position of another scene.
                const unsigned long long other = __ballot(lane < cnt && lane >= start && w_tag != run_tag);
                const uint32_t take = other ? uint32_t(__ffsll(static_cast<long long>(other))) - 1 : cnt;
                const uint32_t k = start + uint32_t(__popcll(free_lanes & ((1ull << lane) - 1ull)));
                const uint32_t k_entry = __shfl(w_entry, int(k & 63u)), k_pos = __shfl(w_pos, int(k & 63u));
                if (idle && k < take) {
                    slot = k_entry & kSlotMask;
                    q.begin(sc, queuedRay<ANY>(ps, slot));
                    if (SEEDED && !ANY) {
                        const uint4 h = ps.hit[slot];
                        q.seed(HitF{__uint_as_float(h.x), h.y, h.z, 0.0f, 0.0f});
                    }
                    id = k_pos;
                }
                next += take;
            }
        }
        // Done when no lane holds a ray and the slice is used up: each pass with
        // every lane idle either starts a run (taking at least one ray) or moves
        // past a window of other kernels' rays, so `next` reaches `end`.
        if (__ballot(id != kNone) == 0 && next >= end) break;
        if (id != kNone) {
            if constexpr (SPHERES) {
                for (uint32_t k = 0; k < kQueueSteps && !q.done; ++k) q.step(sc, stk);
            } else {
                stepQueue(q, sc, stk, next < end);
            }
            if (q.done) {
                finishQuery<ANY>(ps, sc, q, slot, id);
                id = kNone;
            }
        }
    }
}

// Per-block material counts of the keys kTraceQueue wrote (kExtend's counting).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kCountBins(PathState ps)
{
    const uint32_t n = ps.ctr->n_active + pendingGeneration(ps.ctr).n_extra;
    WaveCounts<kBins> wc;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        wc.add(i < n ? ps.key[i] : 0u);
    }
    storeBlockCounts<BLOCK, kBins>(wc, ps.block_count);
}

// ------------------------------------------------------------------ finish

__device__ __forceinline__ unsigned long long toFixed(float v)
{
    if (!(v > 0.0f)) return 0ull;
    return static_cast<unsigned long long>(double(fminf(v, kMaxSampleRadiance)) * kFixedScale + 0.5);
}

// Adds each finished path's radiance to its pixel (reference PixelBuffer::addSample,
// image/pixel_buffer.cpp:29-35) and returns its slot to the free list unless the
// next bounce recycles it.
// Lanes that finish samples of the same pixel are often neighbours (samples of
// a pixel are generated together), so each wave first sums runs of equal
// (job, pixel) keys and only the first lane of a run adds to the pixel: fewer
// and far less contended atomics. Integer sums are exact, so the image is the
// same as with one atomic add per sample.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void kFinish(PathState ps)
{
    Counters* c = ps.ctr;
    const uint32_t n = c->n_term, free0 = c->n_free, n_rec = c->n_recycle;
    const uint32_t lane = __lane_id();
    // Slots pushed on the free stack unsort it.
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > n_rec) c->free_sorted = 0;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        uint32_t j = 0, pix = 0;
        bool counts = false;
        unsigned long long v0 = 0, v1 = 0, v2 = 0, cnt = 0;
        if (i < n) {
            const uint32_t slot = ps.active[ps.termq[i]] & kSlotMask;
            j = __float_as_uint(ps.prev[slot].w);
            counts = !ps.jobs[j].cancelled;
            if (counts) {
                const float4 r = ps.rad[slot];
                v0 = toFixed(r.x);
                v1 = toFixed(r.y);
                v2 = toFixed(r.z);
                cnt = 1;
                pix = ps.pixel[slot];
            }
            if (i >= n_rec) {
                ps.freeq[free0 + (i - n_rec)] = slot;
                ps.slot_free[slot] = 1;
            }
        }
        const unsigned long long key = counts ? (static_cast<unsigned long long>(j) << 32) | pix : ~0ull;
        const unsigned long long up = __shfl_up(key, 1u);
        const bool head = lane == 0 || up != key;
        const unsigned long long heads = __ballot(head);
        if (heads != ~0ull) {
            // Segmented suffix sums: run ids increase along the wave, so equal
            // ids mean the same contiguous run.
            const uint32_t run = uint32_t(__popcll(heads & ((2ull << lane) - 1ull)));
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t r2 = __shfl_down(run, off);
                const unsigned long long a0 = __shfl_down(v0, off), a1 = __shfl_down(v1, off),
                                         a2 = __shfl_down(v2, off), a3 = __shfl_down(cnt, off);
                if (lane + off < 64 && r2 == run) {
                    v0 += a0;
                    v1 += a1;
                    v2 += a2;
                    cnt += a3;
                }
            }
        }
        if (head && cnt) {
            unsigned long long* a = ps.jobs[j].accum + 4 * size_t(pix);
            atomicAdd(a + 0, v0);
            atomicAdd(a + 1, v1);
            atomicAdd(a + 2, v2);
            atomicAdd(a + 3, cnt);
        }
        // Finished samples: once per wave when the wave's paths share a job (the
        // usual case), else once per run, on this block's lane of the job's
        // spread counter.
        const unsigned long long live = __ballot(counts);
        if (live) {
            const size_t dl = size_t(blockIdx.x % kDoneLanes) * kDoneStride;
            const int leader = __ffsll(static_cast<long long>(live)) - 1;
            const uint32_t j0 = __shfl(j, leader);
            if (__ballot(counts && j != j0) == 0) {
                if (lane == uint32_t(leader))
                    atomicAdd(ps.jobs[j0].done + dl, static_cast<unsigned long long>(__popcll(live)));
            } else if (head && cnt) {
                atomicAdd(ps.jobs[j].done + dl, cnt);
            }
        }
    }
}

// Mean radiance per pixel (reference PixelBuffer::average, image/pixel_buffer.cpp:37-45).
__global__ void kResolve(const unsigned long long* accum, uint64_t npx, float* mean)
{
    for (uint64_t p = gridStart(); p < npx; p += gridStride()) {
        const ulonglong2* a = reinterpret_cast<const ulonglong2*>(accum + 4 * p);
        const ulonglong2 x = a[0], y = a[1];
        const double n = double(y.y);
        mean[3 * p + 0] = y.y ? float(double(x.x) / kFixedScale / n) : 0.0f;
        mean[3 * p + 1] = y.y ? float(double(x.y) / kFixedScale / n) : 0.0f;
        mean[3 * p + 2] = y.y ? float(double(y.x) / kFixedScale / n) : 0.0f;
    }
}

__global__ void kReset(PathState ps, unsigned long long total, uint32_t n_segs)
{
    const uint32_t i0 = gridStart();
    // Stacked so the first generation takes slots 0, 1, 2, ... in queue order.
    for (uint32_t i = i0; i < ps.capacity; i += gridStride()) {
        ps.freeq[i] = ps.capacity - 1 - i;
        ps.slot_free[i] = 1;
    }
    if (i0 == 0) {
        Counters* c = ps.ctr;
        *c = Counters{};
        c->n_free = ps.capacity;
        c->free_sorted = 1;
        c->work_total = total;
        c->seg_lo = 0;
        c->seg_hi = n_segs;
    }
}

__global__ void kAddWork(PathState ps, unsigned long long total, uint32_t seg_lo, uint32_t seg_hi)
{
    Counters* c = ps.ctr;
    c->work_total = total;
    c->seg_lo = seg_lo;
    c->seg_hi = seg_hi;
}

// The lane-refill tracing kernel for the shape: over its one scene, or grouped
// by scene (split tracing), with the FP64 sphere test only if a scene it traces
// has spheres, and the rest of a TLAS leaf in a register when that extra stack
// entry would not fit the stack class.
template <int STACK, bool ANY, bool SPHERES, bool REG>
void launchRefill(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    const dim3 g(sh.qgrid), b(64);
    if (sh.refill == kTraceSplit) {
        if (sh.sphere_pass) hipLaunchKernelGGL((kTraceGrouped<STACK, ANY, SPHERES, REG, true>), g, b, 0, s, ps);
        else hipLaunchKernelGGL((kTraceGrouped<STACK, ANY, SPHERES, REG, false>), g, b, 0, s, ps);
    } else if (!SPHERES && sh.sphere_pass)
        hipLaunchKernelGGL((kTraceQueue<STACK, ANY, false, REG, true>), g, b, 0, s, ps, sh.scene);
    else hipLaunchKernelGGL((kTraceQueue<STACK, ANY, SPHERES, REG, false>), g, b, 0, s, ps, sh.scene);
}

// The sphere pass of a kind-split scene, in the stack class of its sphere TLAS
// (blocks of 256 take classes up to 32 entries, as the one-ray-per-lane kernels).
template <int BLOCK, bool ANY>
hipError_t launchSpherePass(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    const dim3 g(sh.grid), b(BLOCK);
    const bool split = sh.refill == kTraceSplit;
    if (sh.sstack <= kShortStack) {
        if (split) hipLaunchKernelGGL((kSpherePassTagged<BLOCK, kShortStack, ANY>), g, b, 0, s, ps);
        else hipLaunchKernelGGL((kSpherePass<BLOCK, kShortStack, ANY>), g, b, 0, s, ps, sh.sphere_scene);
    } else if (sh.sstack <= 32) {
        if (split) hipLaunchKernelGGL((kSpherePassTagged<BLOCK, 32, ANY>), g, b, 0, s, ps);
        else hipLaunchKernelGGL((kSpherePass<BLOCK, 32, ANY>), g, b, 0, s, ps, sh.sphere_scene);
    } else {
        if constexpr (BLOCK < 256) {
            if (sh.sstack > 64) return hipErrorInvalidValue;
            if (split) hipLaunchKernelGGL((kSpherePassTagged<BLOCK, 64, ANY>), g, b, 0, s, ps);
            else hipLaunchKernelGGL((kSpherePass<BLOCK, 64, ANY>), g, b, 0, s, ps, sh.sphere_scene);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

template <int STACK, bool ANY>
void launchTraceQueueS(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    if (sh.q_spheres) {
        if (sh.q_reg) launchRefill<STACK, ANY, true, true>(ps, sh, s);
        else launchRefill<STACK, ANY, true, false>(ps, sh, s);
    } else {
        if (sh.q_reg) launchRefill<STACK, ANY, false, true>(ps, sh, s);
        else launchRefill<STACK, ANY, false, false>(ps, sh, s);
    }
}

// The refill kernels' stack class is the shape's `qstack`; the one-ray-per-lane
// kernels' is the bounce's STACK (the shape's `pstack`).
template <bool ANY>
void launchTraceQueue(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    if (sh.qstack <= kShortStack) launchTraceQueueS<kShortStack, ANY>(ps, sh, s);
    else if (sh.qstack <= kRefillStack24) launchTraceQueueS<kRefillStack24, ANY>(ps, sh, s);
    else if (sh.qstack <= 32) launchTraceQueueS<32, ANY>(ps, sh, s);
    else launchTraceQueueS<64, ANY>(ps, sh, s);
}

template <int BLOCK, int STACK>
hipError_t bounce(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    const uint32_t grid = sh.grid;
    const bool split = sh.refill == kTraceSplit;
    if (sh.free_sort) {
        hipLaunchKernelGGL(kFreeCount, dim3(kFreeGrid), dim3(256), 0, s, ps);
        hipLaunchKernelGGL(kFreeScatter, dim3(kFreeGrid), dim3(256), 0, s, ps);
    }
    if (sh.generate) hipLaunchKernelGGL(kGenerate, dim3(grid), dim3(BLOCK), 0, s, ps);
    if (split) {
        hipLaunchKernelGGL(kTagCount, dim3(kTagGrid), dim3(kTagBlock), 0, s, ps);
        hipLaunchKernelGGL(kTagPlan, dim3(1), dim3(1024), 0, s, ps);
        hipLaunchKernelGGL(kTagScatter, dim3(kTagGrid), dim3(kTagBlock), 0, s, ps);
    }
    if (sh.refill == kTracePlain) {
        hipLaunchKernelGGL((kExtend<BLOCK, STACK, false>), dim3(grid), dim3(BLOCK), 0, s, ps);
    } else {
        if (split) hipLaunchKernelGGL((kExtend<BLOCK, STACK, true>), dim3(grid), dim3(BLOCK), 0, s, ps);
        if (sh.sphere_pass) {
            const hipError_t e = launchSpherePass<BLOCK, false>(ps, sh, s);
            if (e != hipSuccess) return e;
        }
        launchTraceQueue<false>(ps, sh, s);
        hipLaunchKernelGGL((kCountBins<BLOCK>), dim3(grid), dim3(BLOCK), 0, s, ps);
    }
    // One block of 1024 threads, each over ceil(grid / 1024) blocks' counts.
    if (grid <= 1024) hipLaunchKernelGGL(kPlanBins<1>, dim3(1), dim3(1024), 0, s, ps, grid);
    else if (grid <= 2048) hipLaunchKernelGGL(kPlanBins<2>, dim3(1), dim3(1024), 0, s, ps, grid);
    else if (grid <= 4096) hipLaunchKernelGGL(kPlanBins<4>, dim3(1), dim3(1024), 0, s, ps, grid);
    else hipLaunchKernelGGL(kPlanBins<kMaxGrid / 1024>, dim3(1), dim3(1024), 0, s, ps, grid);
    hipLaunchKernelGGL((kScatterBins<BLOCK>), dim3(grid), dim3(BLOCK), 0, s, ps);
    hipLaunchKernelGGL((kShade<BLOCK>), dim3(grid), dim3(BLOCK), 0, s, ps);
    if (split) {
        hipLaunchKernelGGL((kCountShade<BLOCK, true>), dim3(grid), dim3(BLOCK), 0, s, ps);
        hipLaunchKernelGGL((kScatterShade<BLOCK, true>), dim3(grid), dim3(BLOCK), 0, s, ps);
    } else {
        hipLaunchKernelGGL((kCountShade<BLOCK, false>), dim3(grid), dim3(BLOCK), 0, s, ps);
        hipLaunchKernelGGL((kScatterShade<BLOCK, false>), dim3(grid), dim3(BLOCK), 0, s, ps);
    }
    if (sh.refill == kTracePlain)
        hipLaunchKernelGGL((kShadow<BLOCK, STACK, false>), dim3(grid), dim3(BLOCK), 0, s, ps);
    else if (split)
        hipLaunchKernelGGL((kShadow<BLOCK, STACK, true>), dim3(grid), dim3(BLOCK), 0, s, ps);
    if (sh.sphere_pass) {
        const hipError_t e = launchSpherePass<BLOCK, true>(ps, sh, s);
        if (e != hipSuccess) return e;
    }
    if (sh.refill != kTracePlain) launchTraceQueue<true>(ps, sh, s);
    hipLaunchKernelGGL((kFinish<BLOCK>), dim3(grid), dim3(BLOCK), 0, s, ps);
    return hipGetLastError();
}

template <int BLOCK, int STACK>
hipError_t bouncePair(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    PathState swapped = ps;
    swapped.active = ps.active_next;
    swapped.active_next = ps.active;
    const hipError_t e = bounce<BLOCK, STACK>(ps, sh, s);
    if (e != hipSuccess) return e;
    return bounce<BLOCK, STACK>(swapped, sh, s);
}

template <int BLOCK>
hipError_t bounceB(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    if (sh.stack > 64) return hipErrorInvalidValue;
    if (sh.pstack <= kShortStack) return bouncePair<BLOCK, kShortStack>(ps, sh, s);
    if (sh.pstack <= 32) return bouncePair<BLOCK, 32>(ps, sh, s);
    if (sh.pstack <= 64) return bouncePair<BLOCK, 64>(ps, sh, s);
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launchBouncePair(const PathState& ps, const LaunchShape& sh, hipStream_t s)
{
    switch (sh.block) {
    case 64: return bounceB<64>(ps, sh, s);
    case 128: return bounceB<128>(ps, sh, s);
    case 256:
        if (sh.stack > 64) return hipErrorInvalidValue;
        if (sh.pstack <= kShortStack) return bouncePair<256, kShortStack>(ps, sh, s);
        if (sh.pstack <= 32) return bouncePair<256, 32>(ps, sh, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

hipError_t launchReset(const PathState& ps, unsigned long long total, uint32_t n_segs, hipStream_t s)
{
    hipLaunchKernelGGL(kReset, dim3(1024), dim3(256), 0, s, ps, total, n_segs);
    return hipGetLastError();
}

hipError_t launchResolve(const unsigned long long* accum, uint64_t npx, float* mean, hipStream_t s)
{
    if (npx == 0) return hipSuccess;
    const uint64_t want = (npx + 255) / 256;
    const uint64_t blocks = want < 4096 ? want : 4096;
    hipLaunchKernelGGL(kResolve, dim3(uint32_t(blocks)), dim3(256), 0, s, accum, npx, mean);
    return hipGetLastError();
}

namespace {

struct DoneList {
    const unsigned long long* lanes[kDoneListMax];
};

__global__ __launch_bounds__(64) void kSumDone(DoneList l, unsigned long long* out)
{
    // Lane k reads counter k of job blockIdx.x; the wave sums them.
    unsigned long long v = l.lanes[blockIdx.x][size_t(threadIdx.x) * kDoneStride];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}
static_assert(kDoneLanes == 64, "kSumDone gives each of a job's counters one lane of a wave");

} // namespace

hipError_t launchSumDone(unsigned long long* const* done, uint32_t n, unsigned long long* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (n > kDoneListMax) return hipErrorInvalidValue;
    DoneList l{};
    for (uint32_t i = 0; i < n; ++i) l.lanes[i] = done[i];
    hipLaunchKernelGGL(kSumDone, dim3(n), dim3(64), 0, s, l, out);
    return hipGetLastError();
}

hipError_t launchAddWork(const PathState& ps, unsigned long long total, uint32_t seg_lo, uint32_t seg_hi,
                         hipStream_t s)
{
    hipLaunchKernelGGL(kAddWork, dim3(1), dim3(1), 0, s, ps, total, seg_lo, seg_hi);
    return hipGetLastError();
}

} // namespace ab
