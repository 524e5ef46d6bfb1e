// Wavefront engine host code and the rendering entry points of the C ABI.

#include "engine.hpp"

#include "kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

namespace ab {

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

bool finite3(const double* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// Largest max_depth whose draws fit one sample's stride of the pixel stream:
// 2 jitter draws + at most 7 per bounce.
constexpr uint32_t kMaxDepthLimit = uint32_t((1024 - 2) / 7);

// Samples of one pixel generated back to back (a pass of the work order).
constexpr uint32_t kPixelGroup = 16;

// Render accumulators up to this size stay allocated between renders.
constexpr size_t kKeepAccumBytes = size_t(256) << 20;

// Blocks of the grid-stride queue kernels: the configured count, or by default
// enough for one pass over a full queue up to 2048 blocks (8 per CU).
uint32_t gridFor(uint32_t configured, uint32_t capacity, uint32_t block)
{
    if (configured) return std::min(configured, kMaxGrid);
    return std::max<uint32_t>(64, std::min<uint32_t>((capacity + block - 1) / block, 2048));
}

// Traversal stack class of the bounce kernels for a scene stack depth.
uint32_t stackClass(uint32_t depth) { return depth <= kShortStack ? kShortStack : (depth <= 32 ? 32u : 64u); }

// Stack class of the lane-refill kernels for the deepest scene they trace: the
// smallest class that holds the bound and the rest of a TLAS leaf (one entry
// more), else the smallest that holds the bound (the leaf's rest then waits in a
// register).
void chooseRefillStack(LaunchShape& s, uint32_t depth)
{
    const uint32_t classes[] = {kShortStack, kRefillStack24, 32u, 64u};
    s.qstack = 64;
    for (uint32_t k : classes)
        if (depth + 1 <= k) {
            s.qstack = k;
            break;
        }
    s.q_reg = depth + 1 > s.qstack ? 1u : 0u;
}

// Block size and grid for the shape's one-ray-per-lane stack class.
void settleBlocks(LaunchShape& s, const ab_launch_config& cfg, uint32_t capacity)
{
    s.block = cfg.block_size;
    // 256 x 64 stack entries exceed the LDS budget (the one-ray-per-lane kernels'
    // and the sphere pass's classes).
    if (s.block == 256 && std::max(s.pstack, s.sstack) > 32) s.block = 128;
    s.grid = gridFor(cfg.grid_blocks, capacity, s.block);
}

LaunchShape shapeFor(const ab_launch_config& cfg, uint32_t capacity, uint32_t stack)
{
    LaunchShape s{cfg.block_size, cfg.grid_blocks, stackClass(stack)};
    s.pstack = s.stack;
    settleBlocks(s, cfg, capacity);
    return s;
}

// Rays per lane of a full queue in the lane-refill tracing kernels (as in the
// ray-query kernel: slices of 6 rays per lane).
constexpr uint32_t kQueueRaysPerLane = 6;

// Extension and shadow rays are traced with lane refill where incoherent rays
// through triangle BLASes dominate: over the one scene of every job when that
// scene has mesh instances (`one_scene`: the set holds every scene whose paths
// may be in flight); in a queue of several scenes, the rays of the scenes with
// meshes (split), a wave holding one scene at a time (kTraceGrouped). Rays of
// sphere-only scenes keep the one-ray-per-lane kernels, which were faster on
// them (shallow trees of spheres: refill cost the sphere room 50 %).
void chooseTracing(LaunchShape& s, const SceneSet& set, uint32_t capacity, bool one_scene,
                   const ab_launch_config& cfg)
{
    const uint32_t qgrid =
        std::max<uint32_t>((capacity + 64 * kQueueRaysPerLane - 1) / (64 * kQueueRaysPerLane), 1);
    if (one_scene && set.same && set.meshes) {
        s.refill = kTraceOne;
        s.scene = set.first;
        s.q_spheres = set.first.has_spheres ? 1u : 0u;
        chooseRefillStack(s, set.first.stack_depth);
        s.qgrid = qgrid;
        s.pstack = kShortStack; // no ray goes one per lane
        if (set.first.sphere_root != kNoChild && set.first.mesh_root != kNoChild) {
            // Kind-split: the spheres one ray per lane (their FP64 test out of the
            // refill kernel, which then needs fewer registers and diverges less),
            // then the instances seeded with the sphere hits.
            s.sphere_pass = 1;
            s.sphere_scene = set.first;
            s.sphere_scene.tlas_root = set.first.sphere_root;
            s.sphere_scene.stack_depth = set.first.sphere_stack;
            s.sphere_scene.has_meshes = 0;
            s.sstack = stackClass(set.first.sphere_stack);
            s.scene.tlas_root = set.first.mesh_root;
            s.scene.stack_depth = set.first.mesh_stack;
            s.scene.has_spheres = 0;
            s.q_spheres = 0;
            chooseRefillStack(s, set.first.mesh_stack);
        }
        settleBlocks(s, cfg, capacity);
    } else if (set.meshes) {
        s.refill = kTraceSplit;
        s.q_spheres = set.mesh_spheres ? 1u : 0u;
        chooseRefillStack(s, set.deepest_meshes);
        s.qgrid = qgrid;
        if (set.kind_split) {
            // The tagged rays of kind-split scenes: spheres one ray per lane
            // first, then every tagged run seeded (others start from a miss).
            s.sphere_pass = 1;
            s.sstack = stackClass(set.deepest_sphere_pass);
            s.q_spheres = set.unsplit_mesh_spheres ? 1u : 0u;
            chooseRefillStack(s, set.deepest_mesh_pass);
        }
    }
}

bool sameTracing(const LaunchShape& a, const LaunchShape& b)
{
    if (a.refill != b.refill || a.pstack != b.pstack || a.block != b.block || a.free_sort != b.free_sort ||
        a.generate != b.generate)
        return false;
    if (a.refill == kTracePlain) return true;
    return a.qgrid == b.qgrid && a.q_spheres == b.q_spheres && a.q_reg == b.q_reg && a.qstack == b.qstack &&
           a.sphere_pass == b.sphere_pass && a.sstack == b.sstack &&
           std::memcmp(&a.scene, &b.scene, sizeof(DevScene)) == 0 &&
           std::memcmp(&a.sphere_scene, &b.sphere_scene, sizeof(DevScene)) == 0;
}

#define HIP_TRY(expr, what)                                   \
    do {                                                      \
        const hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hipFail(e_, what);       \
    } while (0)

} // namespace

ab_status prepareJob(const JobSpec& spec, DevJob& dj)
{
    const ab_render_params& p = spec.params;
    const ab_camera& c = spec.camera;
    if (!spec.scene) return fail(AB_INVALID_ARGUMENT, "render job has no scene");
    if (p.samples_per_axis == 0)
        return fail(AB_INVALID_ARGUMENT, "samples per axis must be positive, got 0");
    if (p.samples_per_axis > 256)
        return fail(AB_INVALID_ARGUMENT, "samples per axis must be at most 256 (65536 samples per pixel)");
    if (p.width == 0 || p.height == 0 || uint64_t(p.width) * p.height > (uint64_t(1) << 28))
        return fail(AB_INVALID_ARGUMENT, "invalid image size " + std::to_string(p.width) + "x" +
                                             std::to_string(p.height));
    if (p.max_depth > kMaxDepthLimit)
        return fail(AB_INVALID_ARGUMENT, "max_depth must be at most " + std::to_string(kMaxDepthLimit));
    // Camera checks, reference camera/perspective_camera.cpp:24-110.
    const double fov = c.vfov_degrees;
    if (std::isnan(fov) || fov <= 0 || fov >= 180)
        return fail(AB_INVALID_ARGUMENT, "vertical field of view must be in (0, 180) degrees");
    if (!(c.aspect > 0) || !std::isnormal(c.aspect))
        return fail(AB_INVALID_ARGUMENT, "aspect ratio must be a positive normal number");
    if (!finite3(c.look_from) || !finite3(c.look_at) || !finite3(c.up))
        return fail(AB_INVALID_ARGUMENT, "camera position and target must be finite");
    double back[3], up[3];
    for (int k = 0; k < 3; ++k) {
        back[k] = c.look_from[k] - c.look_at[k];
        up[k] = c.up[k];
    }
    const double bl2 = back[0] * back[0] + back[1] * back[1] + back[2] * back[2];
    const double ul2 = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
    if (!std::isnormal(bl2))
        return fail(AB_INVALID_ARGUMENT, "distance from camera position to target must be non-zero and finite");
    if (!std::isnormal(ul2)) return fail(AB_INVALID_ARGUMENT, "camera up direction must be non-zero and finite");
    double w[3], un[3];
    for (int k = 0; k < 3; ++k) {
        w[k] = back[k] / std::sqrt(bl2);
        un[k] = up[k] / std::sqrt(ul2);
    }
    double right[3] = {un[1] * w[2] - un[2] * w[1], un[2] * w[0] - un[0] * w[2], un[0] * w[1] - un[1] * w[0]};
    const double rl2 = right[0] * right[0] + right[1] * right[1] + right[2] * right[2];
    if (rl2 < 1e-12) return fail(AB_INVALID_ARGUMENT, "camera up direction must not be parallel to the view");
    double u[3], v[3];
    for (int k = 0; k < 3; ++k) u[k] = right[k] / std::sqrt(rl2);
    v[0] = w[1] * u[2] - w[2] * u[1];
    v[1] = w[2] * u[0] - w[0] * u[2];
    v[2] = w[0] * u[1] - w[1] * u[0];
    const double hh = std::tan(fov * 3.14159265358979323846 / 360);
    const double hw = c.aspect * hh;
    std::memset(&dj, 0, sizeof(dj));
    for (int k = 0; k < 3; ++k) {
        dj.cam_origin[k] = float(c.look_from[k]);
        dj.cam_forward[k] = float(-w[k]);
        dj.cam_hw[k] = float(u[k] * hw);
        dj.cam_hh[k] = float(v[k] * hh);
    }
    dj.scene = spec.scene->dev;
    dj.width = p.width;
    dj.height = p.height;
    dj.spa = p.samples_per_axis;
    dj.max_depth = p.max_depth;
    dj.roulette_start = p.roulette_start;
    dj.group = std::min<uint32_t>(p.samples_per_axis * p.samples_per_axis, kPixelGroup);
    if (p.estimator > AB_ESTIMATOR_NORMALS)
        return fail(AB_INVALID_ARGUMENT, "unknown estimator " + std::to_string(p.estimator));
    dj.estimator = p.estimator;
    dj.seed = p.seed;
    dj.work_count = uint64_t(p.width) * p.height * p.samples_per_axis * p.samples_per_axis;
    return AB_OK;
}

Engine::~Engine() { freeAll(); }

void Engine::clearTags()
{
    tags_.clear();
    untagged_meshes_ = false;
}

ab_status Engine::tagFor(const DevScene& scene, uint32_t& tag)
{
    tag = 0;
    if (!scene.has_meshes) return AB_OK;
    for (size_t t = 0; t < tags_.size(); ++t)
        if (std::memcmp(&tags_[t], &scene, sizeof(DevScene)) == 0) {
            tag = uint32_t(t + 1);
            return AB_OK;
        }
    if (tags_.size() + 1 >= kMaxTags) return AB_OK; // no tag left: traced one ray per lane
    tags_.push_back(scene);
    tag = uint32_t(tags_.size());
    // The host copy stays in tags_ (never moved: reserved to kMaxTags entries)
    // until the stream has run this copy.
    HIP_TRY(hipMemcpyAsync(d_tags_ + tag, &tags_.back(), sizeof(DevScene), hipMemcpyHostToDevice, stream_),
            "hipMemcpy(scene tag)");
    return AB_OK;
}

void Engine::freeAll()
{
    if (graph_) (void)hipGraphExecDestroy(graph_);
    graph_ = nullptr;
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (arena_) (void)hipFree(arena_);
    if (accum_) (void)hipFree(accum_);
    if (h_counters_) (void)hipHostFree(h_counters_);
    for (hipEvent_t& e : counter_ev_) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (h_segs_) (void)hipHostFree(h_segs_);
    if (stream_) (void)hipStreamDestroy(stream_);
    arena_ = nullptr;
    accum_ = nullptr;
    accum_cap_ = 0;
    h_counters_ = nullptr;
    h_segs_ = nullptr;
    d_segs_ = nullptr;
    stream_ = nullptr;
    capacity_ = max_jobs_ = 0;
    seg_cap_ = seg_lo_ = seg_hi_ = seg_sent_ = 0;
}

ab_status Engine::reserve(uint32_t capacity, uint32_t max_jobs, hipStream_t)
{
    if (capacity == capacity_ && max_jobs <= max_jobs_ && arena_) return AB_OK;
    freeAll();
    HIP_TRY(hipSetDevice(ctx_->device), "hipSetDevice");
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&h_counters_), 3 * sizeof(Counters), hipHostMallocDefault),
            "hipHostMalloc(counters)");
    uint32_t seg_cap = 4096;
    while (seg_cap < max_jobs) seg_cap *= 2;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&h_segs_), sizeof(DevSegment) * seg_cap, hipHostMallocDefault),
            "hipHostMalloc(segments)");
    const size_t C = capacity;
    struct Part {
        void** dst;
        size_t bytes;
    };
    PathState& p = ps_;
    Part parts[] = {{reinterpret_cast<void**>(&p.org), C * 16},      {reinterpret_cast<void**>(&p.dir), C * 16},
                    {reinterpret_cast<void**>(&p.thr), C * 16},      {reinterpret_cast<void**>(&p.rad), C * 16},
                    {reinterpret_cast<void**>(&p.prev), C * 16},     {reinterpret_cast<void**>(&p.rng), C * 8},
                    {reinterpret_cast<void**>(&p.pixel), C * 4},     {reinterpret_cast<void**>(&p.hit), C * 16},
                    {reinterpret_cast<void**>(&p.hit_b2), C * 4},    {reinterpret_cast<void**>(&p.key), C * 4},
                    {reinterpret_cast<void**>(&p.sh_org), C * 16},   {reinterpret_cast<void**>(&p.sh_dir), C * 16},
                    {reinterpret_cast<void**>(&p.sh_c), C * 16},     {reinterpret_cast<void**>(&p.active), C * 4},
                    {reinterpret_cast<void**>(&p.binned), C * 4},    {reinterpret_cast<void**>(&p.shadowq), C * 4},
                    {reinterpret_cast<void**>(&p.termq), C * 4},     {reinterpret_cast<void**>(&p.freeq), C * 4},
                    {reinterpret_cast<void**>(&p.active_next), C * 4},
                    {reinterpret_cast<void**>(&p.bytag), C * 4},
                    {reinterpret_cast<void**>(&p.slot_free), C},
                    {reinterpret_cast<void**>(&p.tag_count), size_t(kMaxTags) * kTagGrid * 4},
                    {reinterpret_cast<void**>(&p.block_count), size_t(kMaxGrid) * 8 * 4},
                    {reinterpret_cast<void**>(&p.block_off), size_t(kMaxGrid) * 8 * 4},
                    {reinterpret_cast<void**>(&p.ctr), sizeof(Counters)},
                    {reinterpret_cast<void**>(&d_jobs_), sizeof(DevJob) * std::max<uint32_t>(max_jobs, 1)},
                    {reinterpret_cast<void**>(&d_tags_), sizeof(DevScene) * kMaxTags},
                    {reinterpret_cast<void**>(&d_segs_), sizeof(DevSegment) * seg_cap}};
    size_t total = 0;
    for (const Part& part : parts) total += align256(part.bytes);
    HIP_TRY(hipMalloc(&arena_, total), "hipMalloc(path state)");
    size_t off = 0;
    for (const Part& part : parts) {
        *part.dst = static_cast<char*>(arena_) + off;
        off += align256(part.bytes);
    }
    p.capacity = capacity;
    p.jobs = d_jobs_;
    p.tag_scenes = d_tags_;
    tags_.clear();
    tags_.reserve(kMaxTags);
    p.segs = d_segs_;
    p.seg_mask = seg_cap - 1;
    seg_cap_ = seg_cap;
    capacity_ = capacity;
    max_jobs_ = std::max<uint32_t>(max_jobs, 1);
    return AB_OK;
}

ab_status Engine::bouncePair(const LaunchShape& shape)
{
    if (use_graph_ && graph_) {
        HIP_TRY(hipGraphLaunch(graph_, stream_), "hipGraphLaunch");
        return AB_OK;
    }
    HIP_TRY(launchBouncePair(ps_, shape, stream_), "bounce launch");
    return AB_OK;
}

namespace {

// Counters as the host uses them: the last bounce's terminations and shadow
// rays, which the device adds at the next bounce, counted now.
Counters settled(Counters c)
{
    c.paths_done += c.n_term;
    c.shadow_rays += c.n_shadow;
    c.n_free += c.n_term - c.n_recycle;
    c.n_term = c.n_shadow = 0;
    return c;
}

} // namespace

ab_status Engine::readCounters(Counters& out)
{
    HIP_TRY(hipMemcpyAsync(h_counters_, ps_.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_),
            "hipMemcpy(counters)");
    HIP_TRY(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    out = settled(*h_counters_);
    return AB_OK;
}

ab_status Engine::enqueueCounters(uint32_t slot)
{
    if (slot > 1) return fail(AB_INVALID_ARGUMENT, "counter slot must be 0 or 1");
    if (!counter_ev_[slot])
        HIP_TRY(hipEventCreateWithFlags(&counter_ev_[slot], hipEventDisableTiming), "hipEventCreate(counters)");
    HIP_TRY(hipMemcpyAsync(h_counters_ + 1 + slot, ps_.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_),
            "hipMemcpy(counters)");
    HIP_TRY(hipEventRecord(counter_ev_[slot], stream_), "hipEventRecord(counters)");
    return AB_OK;
}

ab_status Engine::waitCounters(uint32_t slot, Counters& out)
{
    if (slot > 1 || !counter_ev_[slot]) return fail(AB_INVALID_ARGUMENT, "no counter read in that slot");
    HIP_TRY(hipEventSynchronize(counter_ev_[slot]), "hipEventSynchronize(counters)");
    out = settled(h_counters_[1 + slot]);
    return AB_OK;
}

ab_status Engine::capture(const LaunchShape& shape)
{
    if (!use_graph_) return AB_OK;
    // The graph bakes in the path-state pointers, which change only when
    // reserve() reallocates (and then drops the graph).
    if (graph_ && shape.block == graph_shape_.block && shape.grid == graph_shape_.grid &&
        shape.stack == graph_shape_.stack && shape.pstack == graph_shape_.pstack && sameTracing(shape, graph_shape_))
        return AB_OK;
    if (graph_) {
        // The serving loop enqueues a round before it reads the last one, so the
        // graph may still run: let it finish before it is replaced.
        HIP_TRY(hipStreamSynchronize(stream_), "hipStreamSynchronize(graph)");
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const hipError_t le = launchBouncePair(ps_, shape, stream_);
    const hipError_t ce = hipStreamEndCapture(stream_, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hipFail(le, "bounce capture");
    }
    HIP_TRY(ce, "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        graph_ = nullptr;
        return hipFail(ie, "hipGraphInstantiate");
    }
    graph_shape_ = shape;
    return AB_OK;
}

LaunchShape singleSceneShape(const ab_launch_config& cfg, uint32_t capacity, const DevScene& scene)
{
    LaunchShape shape = shapeFor(cfg, capacity, scene.stack_depth);
    SceneSet set;
    set.add(scene);
    chooseTracing(shape, set, capacity, true, cfg);
    return shape;
}

uint64_t bounceLimit(const DevJob* jobs, size_t n_jobs, uint64_t work_total, uint32_t capacity)
{
    uint64_t max_depth = 0;
    for (size_t i = 0; i < n_jobs; ++i) max_depth = std::max<uint64_t>(max_depth, jobs[i].max_depth);
    return 4 * (max_depth + 2) * (work_total / capacity + 2) + 64;
}

uint32_t firstCheckPairs(const DevJob* jobs, size_t n_jobs)
{
    uint64_t path_bounces = 1; // bounces a path can take (the normals view: one)
    for (size_t i = 0; i < n_jobs; ++i)
        path_bounces = std::max<uint64_t>(path_bounces,
                                          jobs[i].estimator == AB_ESTIMATOR_NORMALS ? 1 : jobs[i].max_depth + 1);
    return uint32_t(std::min<uint64_t>((path_bounces + 2) / 2, 8));
}

ab_status Engine::runUntilDone(const LaunchShape& shape, bool use_graph, Counters& c)
{
    use_graph_ = use_graph;
    // When every sample starts in the first bounce (no free-slot sort), the
    // bounces after the first pair generate nothing: that pair is launched
    // directly and the pairs after it leave the generation kernel out.
    LaunchShape steady = shape;
    bool first_pair = shape.free_sort == 0;
    if (first_pair) steady.generate = 0;
    const ab_status cst = capture(steady);
    if (cst != AB_OK) return cst;
    // Bounces between checks grow geometrically: the queue drains in a tail of
    // short bounces, and each check costs a host round trip.
    const uint64_t limit = bounceLimit(jobs_.data(), jobs_.size(), work_total_, capacity_);
    uint64_t bounces = 0;
    uint32_t pairs_per_check = firstCheckPairs(jobs_.data(), jobs_.size());
    while (true) {
        if (bounces > limit) return fail(AB_DEVICE_ERROR, "render did not finish within its bounce bound");
        bounces += 2 * pairs_per_check;
        for (uint32_t k = 0; k < pairs_per_check; ++k) {
            if (first_pair) {
                HIP_TRY(launchBouncePair(ps_, shape, stream_), "bounce launch");
                first_pair = false;
                continue;
            }
            const ab_status st = bouncePair(steady);
            if (st != AB_OK) return st;
        }
        const ab_status st = readCounters(c);
        if (st != AB_OK) return st;
        if (idle(c)) return AB_OK;
        pairs_per_check = std::min<uint32_t>(pairs_per_check * 2, 32);
    }
}

ab_status Engine::renderBatch(const std::vector<JobSpec>& specs, const std::vector<float*>& out,
                              ab_render_stats* stats, const ab_launch_config& cfg, const std::vector<uint64_t*>* sums)
{
    if (specs.empty()) return AB_OK;
    jobs_.assign(specs.size(), DevJob{});
    uint32_t stack = 0;
    uint64_t total = 0;
    size_t accum_bytes = 0;
    std::vector<JobBuffers> bufs(specs.size());
    for (size_t i = 0; i < specs.size(); ++i) {
        const ab_status st = prepareJob(specs[i], jobs_[i]);
        if (st != AB_OK) return st;
        total += jobs_[i].work_count;
        stack = std::max(stack, jobs_[i].scene.stack_depth);
        bufs[i] = jobBuffers(jobs_[i]);
        accum_bytes += bufs[i].total;
    }
    if (stack > 64) return fail(AB_INVALID_ARGUMENT, "scene BVH too deep for the wavefront engine (stack > 64)");
    const uint32_t capacity = std::max<uint32_t>(cfg.paths_in_flight, 1024);
    ab_status st = reserve(capacity, uint32_t(specs.size()), nullptr);
    if (st != AB_OK) return st;
    // Accumulators up to kKeepAccumBytes stay allocated between renders (an
    // allocation and its free cost about as much host time as a small render's
    // GPU work); larger ones are allocated for this call only.
    struct OneShot {
        void* p{nullptr};
        hipStream_t s{nullptr};
        ~OneShot()
        {
            if (!p) return;
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    } one_shot;
    void* acc = nullptr;
    if (accum_bytes <= kKeepAccumBytes) {
        if (accum_bytes > accum_cap_) {
            // hipFree waits for the device's queued work, so the old buffer is idle when it goes.
            if (accum_) (void)hipFree(accum_);
            accum_ = nullptr;
            accum_cap_ = 0;
            const size_t want = std::min(accum_bytes + accum_bytes / 4, kKeepAccumBytes);
            HIP_TRY(hipMalloc(&accum_, want), "hipMalloc(accumulators)");
            accum_cap_ = want;
        }
        acc = accum_;
    } else {
        HIP_TRY(hipMalloc(&acc, accum_bytes), "hipMalloc(accumulators)");
        one_shot.p = acc;
        one_shot.s = stream_;
    }
    HIP_TRY(hipMemsetAsync(acc, 0, accum_bytes, stream_), "hipMemset(accumulators)");
    size_t off = 0;
    for (size_t i = 0; i < jobs_.size(); ++i) {
        char* base = static_cast<char*>(acc) + off;
        jobs_[i].accum = reinterpret_cast<unsigned long long*>(base);
        jobs_[i].done = reinterpret_cast<unsigned long long*>(base + bufs[i].done_offset);
        off += bufs[i].total;
    }
    work_total_ = total;
    LaunchShape shape = shapeFor(cfg, capacity, stack);
    SceneSet set;
    for (const DevJob& j : jobs_) set.add(j.scene);
    chooseTracing(shape, set, capacity, true, cfg);
    // Every sample starts in the first bounce when they all fit the paths in
    // flight (the free stack starts sorted), so no bounce sorts the free slots.
    shape.free_sort = total > capacity ? 1u : 0u;
    // Split tracing: the jobs of each scene with meshes share a scene tag, and
    // the one-ray-per-lane kernels take the stack class of the untagged scenes.
    clearTags();
    if (shape.refill == kTraceSplit) {
        uint32_t plain_depth = 0;
        for (DevJob& j : jobs_) {
            st = tagFor(j.scene, j.tag);
            if (st != AB_OK) return st;
            if (j.tag == 0) plain_depth = std::max(plain_depth, j.scene.stack_depth);
        }
        shape.pstack = stackClass(plain_depth);
        settleBlocks(shape, cfg, capacity);
    }
    // A render whose samples all fit never has more paths than samples: by
    // default the grid-stride kernels then get about four paths per thread at
    // most (a power of two of blocks, so few graph shapes arise), which shortens
    // the per-bounce scans.
    if (shape.free_sort == 0 && cfg.grid_blocks == 0) {
        const uint64_t want = (total + 4ull * shape.block - 1) / (4ull * shape.block);
        uint32_t g = 64;
        while (g < want && g < shape.grid) g *= 2;
        shape.grid = std::min(g, shape.grid);
    }
    HIP_TRY(hipMemcpyAsync(d_jobs_, jobs_.data(), sizeof(DevJob) * jobs_.size(), hipMemcpyHostToDevice, stream_),
            "hipMemcpy(jobs)");
    // One segment per job, in job order (the ring holds at least max_jobs entries).
    HIP_TRY(hipStreamSynchronize(stream_), "hipStreamSynchronize(segments)"); // the mirror is free to rewrite
    uint64_t begin = 0;
    for (size_t i = 0; i < jobs_.size(); ++i) {
        h_segs_[i] = DevSegment{begin, 0, uint32_t(i), 0};
        begin += jobs_[i].work_count;
    }
    seg_lo_ = 0;
    seg_hi_ = seg_sent_ = uint32_t(jobs_.size());
    HIP_TRY(hipMemcpyAsync(d_segs_, h_segs_, sizeof(DevSegment) * jobs_.size(), hipMemcpyHostToDevice, stream_),
            "hipMemcpy(segments)");
    // The timing events, destroyed on every way out of this function.
    struct Events {
        hipEvent_t e[2]{nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e[0]), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ev.e[1]), "hipEventCreate");
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    HIP_TRY(launchReset(ps_, total, uint32_t(jobs_.size()), stream_), "reset launch");
    HIP_TRY(hipEventRecord(e0, stream_), "hipEventRecord");
    Counters c{};
    st = runUntilDone(shape, cfg.use_graph != 0, c);
    (void)hipEventRecord(e1, stream_);
    std::vector<uint64_t> words; // sums mode: one job's accumulators, then added to the caller's
    for (size_t i = 0; st == AB_OK && i < jobs_.size(); ++i) {
        const DevJob& j = jobs_[i];
        const size_t npx = size_t(j.width) * j.height;
        if (sums) {
            words.resize(npx * 4);
            hipError_t e = hipMemcpyAsync(words.data(), j.accum, npx * 32, hipMemcpyDeviceToHost, stream_);
            if (e == hipSuccess) e = hipStreamSynchronize(stream_);
            if (e != hipSuccess) {
                st = hipFail(e, "sums readback");
                break;
            }
            uint64_t* dst = (*sums)[i];
            for (size_t k = 0; k < npx * 4; ++k) dst[k] += words[k];
            continue;
        }
        float* mean = reinterpret_cast<float*>(reinterpret_cast<char*>(j.accum) + bufs[i].mean_offset);
        hipError_t e = launchResolve(j.accum, npx, mean, stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(out[i], mean, npx * 12, hipMemcpyDeviceToHost, stream_);
        if (e != hipSuccess) st = hipFail(e, "image readback");
    }
    if (st == AB_OK) {
        const hipError_t e = hipStreamSynchronize(stream_);
        if (e != hipSuccess) st = hipFail(e, "hipStreamSynchronize(images)");
    }
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    // On success the image copies above synchronised the stream; after a failure
    // the next render must not find work of this one still queued.
    if (st != AB_OK) (void)hipStreamSynchronize(stream_);
    if (st == AB_OK && stats) {
        stats->camera_rays = c.camera_rays;
        stats->extension_rays = c.extend_rays - c.camera_rays;
        stats->shadow_rays = c.shadow_rays;
        stats->paths = c.paths_done;
        stats->gpu_ms = ms;
    }
    return st;
}

ab_status Engine::beginStream(const ab_launch_config& cfg, uint32_t max_jobs)
{
    const uint32_t capacity = std::max<uint32_t>(cfg.paths_in_flight, 1024);
    ab_status st = reserve(capacity, max_jobs, nullptr);
    if (st != AB_OK) return st;
    jobs_.clear();
    jobs_.reserve(max_jobs_);
    clearTags();
    work_total_ = 0;
    seg_lo_ = seg_hi_ = seg_sent_ = 0;
    stream_cfg_ = cfg;
    use_graph_ = cfg.use_graph != 0;
    HIP_TRY(launchReset(ps_, 0, 0, stream_), "reset launch");
    return setStreamStack(kShortStack);
}

ab_status Engine::traceScenes(const SceneSet* live)
{
    // With the scenes in flight known, the one-ray-per-lane kernels take the stack
    // class of the scenes they trace: all of them when none has meshes, else the
    // untagged ones (sphere-only scenes, and every scene once a mesh scene went
    // without a tag). A deep mesh scene in the mix then no longer costs the sphere
    // scenes' rays LDS and occupancy.
    uint32_t stack = shape_.stack;
    if (live) {
        const uint32_t depth = !live->meshes || untagged_meshes_
                                   ? std::max(live->deepest_plain, live->deepest_meshes)
                                   : live->deepest_plain;
        stack = stackClass(depth);
    }
    LaunchShape next = shapeFor(stream_cfg_, capacity_, stack);
    chooseTracing(next, live ? *live : table_set_, capacity_, live != nullptr, stream_cfg_);
    if (next.stack == shape_.stack && sameTracing(next, shape_)) return AB_OK;
    shape_ = next;
    return capture(shape_);
}

ab_status Engine::setStreamStack(uint32_t stack)
{
    shape_ = shapeFor(stream_cfg_, capacity_, stack);
    chooseTracing(shape_, table_set_, capacity_, false, stream_cfg_);
    return capture(shape_);
}

ab_status Engine::appendJob(const DevJob& prepared, unsigned long long* accum, unsigned long long* done,
                            uint32_t* index)
{
    if (tableFull()) return fail(AB_REJECTED_BUDGET, "job table full");
    if (prepared.scene.stack_depth > 64)
        return fail(AB_INVALID_ARGUMENT, "scene BVH too deep for the wavefront engine (stack > 64)");
    // The bounce kernels follow the scenes in the table (the deepest one's stack
    // class; lane refill once one has meshes); queues do not depend on them, so
    // they can change between bounces.
    const SceneSet before = table_set_;
    table_set_.add(prepared.scene);
    const uint32_t stack = std::max(shape_.stack, stackClass(prepared.scene.stack_depth));
    LaunchShape next = shapeFor(stream_cfg_, capacity_, stack);
    chooseTracing(next, table_set_, capacity_, false, stream_cfg_);
    if (next.stack != shape_.stack || !sameTracing(next, shape_)) {
        const ab_status st = setStreamStack(stack);
        if (st != AB_OK) {
            table_set_ = before;
            return st;
        }
    }
    DevJob dj = prepared;
    dj.accum = accum;
    dj.done = done;
    // Every job of a scene with meshes gets that scene's tag while tags last, so
    // whichever way the bounces trace (they switch with the scenes in flight), its
    // paths' queue entries name their scene.
    const ab_status tst = tagFor(dj.scene, dj.tag);
    if (tst != AB_OK) return tst;
    if (dj.scene.has_meshes && dj.tag == 0) untagged_meshes_ = true;
    const uint32_t idx = uint32_t(jobs_.size());
    jobs_.push_back(dj);
    HIP_TRY(hipMemcpyAsync(d_jobs_ + idx, &jobs_[idx], sizeof(DevJob), hipMemcpyHostToDevice, stream_),
            "hipMemcpy(job)");
    if (index) *index = idx;
    return AB_OK;
}

bool Engine::exposeWork(uint32_t index, uint64_t local, uint64_t count)
{
    if (count == 0) return true;
    if (seg_hi_ - seg_lo_ >= seg_cap_) return false;
    h_segs_[seg_hi_ & (seg_cap_ - 1)] = DevSegment{work_total_, local, index, 0};
    ++seg_hi_;
    work_total_ += count;
    return true;
}

ab_status Engine::publishWork(unsigned long long work_next)
{
    // Segment seg_lo_ + 1 beginning at or before work_next means seg_lo_ is spent.
    while (seg_hi_ - seg_lo_ > 1 && h_segs_[(seg_lo_ + 1) & (seg_cap_ - 1)].begin <= work_next) ++seg_lo_;
    if (seg_sent_ == seg_hi_) return AB_OK;
    // Copy the new entries, in two runs where they wrap around the ring. The
    // mirror entries stay untouched until the stream has run these copies: an
    // entry is rewritten only after it retires, which needs a later counter read.
    const uint32_t mask = seg_cap_ - 1;
    uint32_t from = seg_sent_;
    while (from != seg_hi_) {
        const uint32_t slot = from & mask;
        const uint32_t run = std::min<uint32_t>(seg_hi_ - from, seg_cap_ - slot);
        HIP_TRY(hipMemcpyAsync(d_segs_ + slot, h_segs_ + slot, sizeof(DevSegment) * run, hipMemcpyHostToDevice,
                               stream_),
                "hipMemcpy(segments)");
        from += run;
    }
    seg_sent_ = seg_hi_;
    HIP_TRY(launchAddWork(ps_, work_total_, seg_lo_, seg_hi_, stream_), "add-work launch");
    return AB_OK;
}

ab_status Engine::cancelJob(uint32_t index)
{
    if (index >= jobs_.size()) return fail(AB_NOT_FOUND, "no such engine job");
    jobs_[index].cancelled = 1;
    HIP_TRY(hipMemcpyAsync(&d_jobs_[index].cancelled, &jobs_[index].cancelled, sizeof(uint32_t),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(cancel)");
    return AB_OK;
}

ab_status Engine::runBounces(uint32_t count)
{
    for (uint32_t k = 0; k < count; k += 2) {
        const ab_status st = bouncePair(shape_);
        if (st != AB_OK) return st;
    }
    return AB_OK;
}

ab_status Engine::resetTable()
{
    jobs_.clear();
    clearTags();
    work_total_ = 0;
    seg_lo_ = seg_hi_ = seg_sent_ = 0;
    HIP_TRY(launchReset(ps_, 0, 0, stream_), "reset launch");
    table_set_ = SceneSet{};
    return shape_.stack > kShortStack || shape_.refill != kTracePlain ? setStreamStack(kShortStack) : AB_OK;
}

Engine& contextEngine(ab_context* ctx)
{
    if (!ctx->engine) ctx->engine = new Engine(ctx);
    return *ctx->engine;
}

} // namespace ab

using namespace ab;

extern "C" {

ab_status ab_render(ab_scene* scene, const ab_camera* camera, const ab_render_params* params, float* rgb_mean,
                    ab_render_stats* stats)
{
    if (!scene || !camera || !params || !rgb_mean) return fail(AB_INVALID_ARGUMENT, "ab_render: null argument");
    ab_context* ctx = scene->ctx;
    std::lock_guard<std::mutex> lock(ctx->render_mu);
    ab_launch_config cfg;
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        cfg = ctx->config;
    }
    JobSpec js{scene, *camera, *params};
    return contextEngine(ctx).renderBatch({js}, {rgb_mean}, stats, cfg);
}

ab_status ab_render_sums(ab_scene* scene, const ab_camera* camera, const ab_render_params* params, uint64_t* sums,
                         ab_render_stats* stats)
{
    if (!scene || !camera || !params || !sums) return fail(AB_INVALID_ARGUMENT, "ab_render_sums: null argument");
    ab_context* ctx = scene->ctx;
    std::lock_guard<std::mutex> lock(ctx->render_mu);
    ab_launch_config cfg;
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        cfg = ctx->config;
    }
    JobSpec js{scene, *camera, *params};
    const std::vector<uint64_t*> dst{sums};
    return contextEngine(ctx).renderBatch({js}, {nullptr}, stats, cfg, &dst);
}

ab_status ab_render_batch(ab_context* ctx, const ab_job_desc* jobs, uint32_t n_jobs, float* const* rgb_means,
                          ab_render_stats* stats)
{
    if (!ctx || (n_jobs && (!jobs || !rgb_means))) return fail(AB_INVALID_ARGUMENT, "ab_render_batch: null argument");
    std::vector<JobSpec> specs;
    std::vector<float*> outs;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        if (!rgb_means[i]) return fail(AB_INVALID_ARGUMENT, "ab_render_batch: null output");
        if (jobs[i].scene && jobs[i].scene->ctx != ctx)
            return fail(AB_INVALID_ARGUMENT, "ab_render_batch: scene belongs to another context");
        specs.push_back({jobs[i].scene, jobs[i].camera, jobs[i].params});
        outs.push_back(rgb_means[i]);
    }
    std::lock_guard<std::mutex> lock(ctx->render_mu);
    ab_launch_config cfg;
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        cfg = ctx->config;
    }
    return contextEngine(ctx).renderBatch(specs, outs, stats, cfg);
}

ab_status ab_autotune(ab_context* ctx, const ab_job_desc* jobs, uint32_t n_jobs, uint32_t mode,
                      ab_launch_config* best, double* best_ms)
{
    if (!ctx || !jobs || n_jobs == 0) return fail(AB_INVALID_ARGUMENT, "ab_autotune: no jobs");
    if (mode != AB_TUNE_PLAIN && mode != AB_TUNE_BATCH)
        return fail(AB_INVALID_ARGUMENT, "ab_autotune: mode must be AB_TUNE_PLAIN or AB_TUNE_BATCH");
    uint64_t samples = 0;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        if (!jobs[i].scene || jobs[i].scene->ctx != ctx)
            return fail(AB_INVALID_ARGUMENT, "ab_autotune: job without a scene of this context");
        const uint64_t spa = jobs[i].params.samples_per_axis;
        samples += uint64_t(jobs[i].params.width) * jobs[i].params.height * spa * spa;
    }
    // Trial size: every job keeps its scene, camera (aspect ratio included) and
    // samples per pixel and renders a band of its rows, the same share of each
    // job's rows, so one trial takes about kTuneSamples camera samples. The queue
    // then holds jobs of the same sizes relative to each other, and pixels of the
    // same sample counts, as the real run; cutting samples per pixel instead would
    // leave every job so small that the whole list sits in the queue at once.
    // Trials of 2^26 samples on the 64-job batched workload (2.4 G samples)
    // ranked 4M paths in flight first, against 2M in the full run; from 2^28 the
    // ranking matched.
    constexpr uint64_t kTuneSamples = uint64_t(1) << 28;
    const double share = samples > kTuneSamples ? double(kTuneSamples) / double(samples) : 1.0;
    std::vector<JobSpec> specs;
    std::vector<std::vector<float>> images(n_jobs);
    std::vector<float*> outs;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        JobSpec s{jobs[i].scene, jobs[i].camera, jobs[i].params};
        s.params.height = std::max<uint32_t>(1, uint32_t(std::ceil(double(s.params.height) * share)));
        DevJob check{};
        const ab_status st = prepareJob(s, check);
        if (st != AB_OK) return st;
        specs.push_back(s);
        images[i].resize(size_t(s.params.width) * s.params.height * 3);
        outs.push_back(images[i].data());
    }
    std::lock_guard<std::mutex> lock(ctx->render_mu);
    ab_launch_config base;
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        base = ctx->config;
    }
    Engine& engine = contextEngine(ctx);
    auto trial = [&](const ab_launch_config& cfg, double& ms) -> ab_status {
        const auto t0 = std::chrono::steady_clock::now();
        if (mode == AB_TUNE_BATCH) {
            const ab_status st = engine.renderBatch(specs, outs, nullptr, cfg);
            if (st != AB_OK) return st;
        } else {
            for (size_t i = 0; i < specs.size(); ++i) {
                const ab_status st = engine.renderBatch({specs[i]}, {outs[i]}, nullptr, cfg);
                if (st != AB_OK) return st;
            }
        }
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return AB_OK;
    };
    const uint32_t blocks[] = {128, 256};
    const uint32_t paths[] = {1u << 20, 1u << 21, 1u << 22};
    const uint32_t grids[] = {0, 4096};
    ab_launch_config winner = base;
    double winner_ms = 1e300;
    for (uint32_t b : blocks)
        for (uint32_t p : paths)
            for (uint32_t g : grids) {
                const ab_launch_config cfg{b, p, g, base.use_graph};
                // The first render under a configuration allocates its path state
                // and captures its bounce graph: a one-row render of the first job
                // does that before the timed trial.
                JobSpec warm = specs[0];
                warm.params.height = 1;
                float* row = images[0].data();
                ab_status st = engine.renderBatch({warm}, {row}, nullptr, cfg);
                double ms = 0;
                if (st == AB_OK) st = trial(cfg, ms);
                if (st != AB_OK) return st;
                if (ms < winner_ms) {
                    winner_ms = ms;
                    winner = cfg;
                }
            }
    {
        std::lock_guard<std::mutex> l2(ctx->mu);
        ctx->config = winner;
    }
    if (best) *best = winner;
    if (best_ms) *best_ms = winner_ms;
    return AB_OK;
}

ab_status ab_set_launch_config(ab_context* ctx, const ab_launch_config* c)
{
    if (!ctx || !c) return fail(AB_INVALID_ARGUMENT, "ab_set_launch_config: null argument");
    if (c->block_size != 64 && c->block_size != 128 && c->block_size != 256)
        return fail(AB_INVALID_ARGUMENT, "block_size must be 64, 128 or 256");
    if (c->paths_in_flight < 1024 || c->paths_in_flight > (1u << 26))
        return fail(AB_INVALID_ARGUMENT, "paths_in_flight must be in [1024, 2^26]");
    if (c->grid_blocks > kMaxGrid)
        return fail(AB_INVALID_ARGUMENT, "grid_blocks must be at most " + std::to_string(kMaxGrid));
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->config = *c;
    return AB_OK;
}

ab_status ab_get_launch_config(const ab_context* ctx, ab_launch_config* c)
{
    if (!ctx || !c) return fail(AB_INVALID_ARGUMENT, "ab_get_launch_config: null argument");
    std::lock_guard<std::mutex> lock(const_cast<ab_context*>(ctx)->mu);
    *c = ctx->config;
    return AB_OK;
}

} // extern "C"
