// ab_render_static: the wavefront path tracer of ab_render for callers that
// cannot allocate (tight serving loops, harnesses that forbid hipMalloc). Path
// state, queues, counters, the job entry and the per-pixel accumulators live in
// one file-scope device array, `scratch`, of kScratchBytes on every GPU the
// library runs on; the image goes to a device buffer the caller owns, on the
// caller's stream. No hipMalloc, hipFree, host allocation of pinned memory,
// graph instantiation, event or stream creation happens in the call.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two host threads, or two streams without ordering between them) race and
// corrupt each other's images. Each call clears what it uses before it starts.

#include "engine.hpp"
#include "kernels.hpp"
#include "wavefront.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

namespace ab {

namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Path slots resident at once, and the largest image (1920 x 1080 fits).
constexpr uint32_t kStaticSlots = 1u << 17;
constexpr uint64_t kStaticPixels = uint64_t(1) << 21;

// Bytes of each part of the array, in order.
constexpr size_t kSlotBytes16 = size_t(kStaticSlots) * 16;
constexpr size_t kSlotBytes8 = size_t(kStaticSlots) * 8;
constexpr size_t kSlotBytes4 = size_t(kStaticSlots) * 4;
constexpr size_t kBlockBytes = size_t(kMaxGrid) * 8 * 4;
// org dir thr rad prev hit sh_org sh_dir sh_c (16 B), rng (8 B), pixel hit_b2 key
// active active_next binned shadowq termq freeq (4 B), the slots' free flags
// (1 B), per-block counts and offsets, counters, the job, the segment, the
// finished-sample counter and the accumulators (32 B per pixel).
constexpr size_t kPathBytes = 9 * align256(kSlotBytes16) + align256(kSlotBytes8) + 9 * align256(kSlotBytes4) +
                              align256(kStaticSlots);
constexpr size_t kFixedBytes = kPathBytes + 2 * align256(kBlockBytes) + align256(sizeof(Counters)) +
                               align256(sizeof(DevJob)) + align256(sizeof(DevSegment)) + align256(kDoneBytes);
constexpr size_t kScratchBytes = kFixedBytes + size_t(kStaticPixels) * 32;
static_assert(kScratchBytes <= (size_t(96) << 20), "the no-allocation renderer's array is at most 96 MiB");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each GPU (each holds its own copy), looked up
// once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

struct StaticLayout {
    PathState ps{};
    DevJob* job{nullptr};
    DevSegment* seg{nullptr};
    unsigned long long* done{nullptr};
    unsigned long long* accum{nullptr};
};

StaticLayout carve(void* base)
{
    char* p = static_cast<char*>(base);
    const auto take = [&p](size_t bytes) {
        char* r = p;
        p += align256(bytes);
        return r;
    };
    StaticLayout L;
    PathState& s = L.ps;
    s.org = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.dir = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.thr = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.rad = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.prev = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.hit = reinterpret_cast<uint4*>(take(kSlotBytes16));
    s.sh_org = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.sh_dir = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.sh_c = reinterpret_cast<float4*>(take(kSlotBytes16));
    s.rng = reinterpret_cast<uint64_t*>(take(kSlotBytes8));
    s.pixel = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.hit_b2 = reinterpret_cast<float*>(take(kSlotBytes4));
    s.key = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.active = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.active_next = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.binned = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.shadowq = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.termq = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.freeq = reinterpret_cast<uint32_t*>(take(kSlotBytes4));
    s.slot_free = reinterpret_cast<uint8_t*>(take(kStaticSlots));
    s.block_count = reinterpret_cast<uint32_t*>(take(kBlockBytes));
    s.block_off = reinterpret_cast<uint32_t*>(take(kBlockBytes));
    s.ctr = reinterpret_cast<Counters*>(take(sizeof(Counters)));
    // bytag and tag_count stay null: a render of one scene never traces split.
    L.job = reinterpret_cast<DevJob*>(take(sizeof(DevJob)));
    L.seg = reinterpret_cast<DevSegment*>(take(sizeof(DevSegment)));
    L.done = reinterpret_cast<unsigned long long*>(take(kDoneBytes));
    L.accum = reinterpret_cast<unsigned long long*>(p);
    s.jobs = L.job;
    s.segs = L.seg;
    s.capacity = kStaticSlots;
    s.seg_mask = 0;
    return L;
}

// Clears what a render reads before writing it: accumulators and finished-sample
// counter to zero, every slot free, the counters of a fresh engine; writes the
// job entry and its one segment. (A fill kernel rather than hipMemsetAsync, so a
// caller's captured stream records no memset node.)
__global__ void kStaticInit(PathState ps, DevJob job, unsigned long long* accum, unsigned long long* done,
                            uint64_t accum_words)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t i0 = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < accum_words; i += stride) accum[i] = 0ull;
    for (uint64_t i = i0; i < kDoneBytes / 8; i += stride) done[i] = 0ull;
    for (uint64_t i = i0; i < ps.capacity; i += stride) {
        ps.freeq[i] = uint32_t(ps.capacity - 1 - i);
        ps.slot_free[i] = 1;
    }
    if (i0 == 0) {
        *const_cast<DevJob*>(ps.jobs) = job;
        *const_cast<DevSegment*>(ps.segs) = DevSegment{0, 0, 0, 0};
        Counters c{};
        c.n_free = ps.capacity;
        c.free_sorted = 1;
        c.work_total = job.work_count;
        c.seg_lo = 0;
        c.seg_hi = 1;
        *ps.ctr = c;
    }
}

#define HIP_TRY(expr, what)                                   \
    do {                                                      \
        const hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hipFail(e_, what);       \
    } while (0)

} // namespace

// Looks the array up on `device`, which must be current: once per device, at
// context creation. The status is kept for the entry point to report.
void staticScratchSetup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_status[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
}

} // namespace ab

using namespace ab;

extern "C" ab_status ab_render_static(ab_scene* scene, const ab_camera* camera, const ab_render_params* params,
                                      float* d_rgb_mean, void* stream)
{
    if (!scene || !camera || !params || !d_rgb_mean)
        return fail(AB_INVALID_ARGUMENT, "ab_render_static: null argument");
    ab_context* ctx = scene->ctx;
    DevJob job{};
    const ab_status pst = prepareJob(JobSpec{scene, *camera, *params}, job);
    if (pst != AB_OK) return pst;
    const uint64_t npx = uint64_t(params->width) * params->height;
    if (npx > kStaticPixels)
        return fail(AB_INVALID_ARGUMENT, "ab_render_static: an image of " + std::to_string(npx) +
                                             " pixels exceeds the no-allocation renderer's " +
                                             std::to_string(kStaticPixels) + "-pixel limit");
    if (job.scene.stack_depth > 64)
        return fail(AB_INVALID_ARGUMENT, "scene BVH too deep for the wavefront engine (stack > 64)");
    const int dev = ctx->device;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (dev < 0 || dev >= kMaxDevices || !g_done[dev])
            return fail(AB_DEVICE_ERROR, "ab_render_static: no-allocation array not set up on device " +
                                             std::to_string(dev));
        if (g_status[dev] != hipSuccess) return hipFail(g_status[dev], "hipGetSymbolAddress(scratch)");
        base = g_ptr[dev];
    }
    HIP_TRY(hipSetDevice(dev), "hipSetDevice");
    ab_launch_config cfg;
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        cfg = ctx->config;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const StaticLayout L = carve(base);
    job.accum = L.accum;
    job.done = L.done;
    const LaunchShape shape = singleSceneShape(cfg, kStaticSlots, job.scene);
    hipLaunchKernelGGL(kStaticInit, dim3(1024), dim3(256), 0, s, L.ps, job, L.accum, L.done, npx * 4);
    HIP_TRY(hipGetLastError(), "init launch");
    // The engine's bounce loop with direct launches: pairs of bounces between
    // counter checks, growing geometrically, within the bounce bound.
    const uint64_t limit = bounceLimit(&job, 1, job.work_count, kStaticSlots);
    uint32_t pairs = firstCheckPairs(&job, 1);
    uint64_t bounces = 0;
    Counters c{};
    while (true) {
        if (bounces > limit) return fail(AB_DEVICE_ERROR, "render did not finish within its bounce bound");
        bounces += 2 * pairs;
        for (uint32_t k = 0; k < pairs; ++k) HIP_TRY(launchBouncePair(L.ps, shape, s), "bounce launch");
        HIP_TRY(hipMemcpyAsync(&c, L.ps.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), "hipMemcpy(counters)");
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (c.work_next == c.work_total && c.n_active == 0) break;
        pairs = pairs * 2 < 32 ? pairs * 2 : 32;
    }
    HIP_TRY(launchResolve(L.accum, npx, d_rgb_mean, s), "resolve launch");
    return AB_OK;
}
