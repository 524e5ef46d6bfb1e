// Host side of the wavefront engine: device memory for path state and queues,
// the job table, the bounce loop (optionally replayed as a hipGraph).
#pragma once

#include "context.hpp"
#include "wavefront.hpp"

#include <cstring>
#include <vector>

namespace ab {

// A validated job with its camera frame, ready for the device table.
struct JobSpec {
    ab_scene* scene{nullptr};
    ab_camera camera{};
    ab_render_params params{};
};

// Checks params and camera as the reference does (render_image.cpp:60-67,
// perspective_camera.cpp:24-110) and fills the FP32 camera frame of `dj`.
ab_status prepareJob(const JobSpec& spec, DevJob& dj);

// Device memory of one job, in one allocation: the per-pixel accumulators, the
// finished-sample counter and the resolved mean image (float RGB per pixel).
struct JobBuffers {
    size_t done_offset;
    size_t mean_offset;
    size_t total;
};

inline JobBuffers jobBuffers(const DevJob& j)
{
    const auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t npx = size_t(j.width) * j.height;
    JobBuffers b;
    b.done_offset = align(npx * 32);
    b.mean_offset = b.done_offset + kDoneBytes;
    b.total = b.mean_offset + align(npx * 12);
    return b;
}

// What the bounce kernels must handle of a set of jobs' scenes.
struct SceneSet {
    bool meshes{false}, same{true}, empty{true};
    bool mesh_spheres{false};  // a scene with meshes also has spheres
    uint32_t deepest_meshes{0}; // largest stack bound of the scenes with meshes
    uint32_t deepest_plain{0};  // ... of the scenes without
    // Scenes with meshes traced in two passes (kind-split, DevScene::sphere_root):
    bool kind_split{false};            // any
    uint32_t deepest_sphere_pass{0};   // their sphere TLASes' largest stack bound
    uint32_t deepest_mesh_pass{0};     // the refill pass's: their instance TLASes', the others' whole
    bool unsplit_mesh_spheres{false};  // a scene with meshes and spheres traced whole
    DevScene first{};
    void add(const DevScene& s)
    {
        if (empty) first = s;
        else if (std::memcmp(&s, &first, sizeof(DevScene)) != 0) same = false;
        empty = false;
        if (!s.has_meshes) deepest_plain = deepest_plain > s.stack_depth ? deepest_plain : s.stack_depth;
        if (s.has_meshes) {
            meshes = true;
            mesh_spheres = mesh_spheres || s.has_spheres;
            deepest_meshes = deepest_meshes > s.stack_depth ? deepest_meshes : s.stack_depth;
            uint32_t pass = s.stack_depth;
            if (s.sphere_root != kNoChild && s.mesh_root != kNoChild) {
                kind_split = true;
                deepest_sphere_pass = deepest_sphere_pass > s.sphere_stack ? deepest_sphere_pass : s.sphere_stack;
                pass = s.mesh_stack;
            } else if (s.has_spheres) {
                unsplit_mesh_spheres = true;
            }
            deepest_mesh_pass = deepest_mesh_pass > pass ? deepest_mesh_pass : pass;
        }
    }
};

class Engine {
public:
    explicit Engine(ab_context* ctx) : ctx_(ctx) {}
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Allocates path state for `capacity` slots and room for `max_jobs` jobs.
    ab_status reserve(uint32_t capacity, uint32_t max_jobs, hipStream_t stream);

    // Renders every job to completion together; out[i] receives job i's
    // per-pixel mean (width*height*3 floats), or, when `sums` is given, job i's
    // per-pixel sums are added to sums[i] (width*height*4 words, the wire format
    // of ab_job_poll_sums) and `out` is unused.
    ab_status renderBatch(const std::vector<JobSpec>& jobs, const std::vector<float*>& out, ab_render_stats* stats,
                          const ab_launch_config& cfg, const std::vector<uint64_t*>* sums = nullptr);

    // Serving interface: jobs are appended while the engine runs. The job table
    // holds `max_jobs` entries and is emptied by resetTable() once idle. A job's
    // samples are generated only as the caller exposes them, in segments
    // (exposeWork, then publishWork), so the caller decides how jobs interleave.
    ab_status beginStream(const ab_launch_config& cfg, uint32_t max_jobs);
    ab_status appendJob(const DevJob& prepared, unsigned long long* accum, unsigned long long* done,
                        uint32_t* index);
    // Queues work indices [local, local + count) of job `index` after all work
    // exposed so far. False when the segment ring is full (retry after publishWork).
    bool exposeWork(uint32_t index, uint64_t local, uint64_t count);
    // Retires segments generated before `work_next` (the device's counter read
    // while the stream was idle) and hands the queued segments to the device.
    ab_status publishWork(unsigned long long work_next);
    unsigned long long workTotal() const { return work_total_; }
    ab_status cancelJob(uint32_t index);
    ab_status runBounces(uint32_t count); // rounded up to an even count
    // Serving: the scenes whose paths may be in flight, when known exactly (every
    // job with paths left among them), else null for every scene of the table.
    // Their being one scene with meshes lets the bounces trace over that scene
    // alone (as a render of one job does); recaptures the bounces if that changes.
    ab_status traceScenes(const SceneSet* live);
    ab_status resetTable();
    bool tableFull() const { return jobs_.size() >= max_jobs_; }
    size_t tableSize() const { return jobs_.size(); }
    ab_status readCounters(Counters& out);
    // The same read without waiting: a copy into pinned slot `slot` (0 or 1) and
    // an event after it, enqueued behind the work so far; waitCounters waits for
    // that event and returns the slot's counters. The serving loop reads one
    // round's counters while the next round runs.
    ab_status enqueueCounters(uint32_t slot);
    ab_status waitCounters(uint32_t slot, Counters& out);
    bool idle(const Counters& c) const { return c.work_next == c.work_total && c.n_active == 0; }
    uint32_t capacity() const { return capacity_; }
    hipStream_t stream() const { return stream_; }

private:
    ab_status runUntilDone(const LaunchShape& shape, bool use_graph, Counters& final_counters);
    ab_status bouncePair(const LaunchShape& shape);
    // (Re)captures the bounce-pair graph for `shape` unless it is current.
    ab_status capture(const LaunchShape& shape);
    // Serving: switches the bounce kernels to the stack class of `stack`.
    ab_status setStreamStack(uint32_t stack);
    // The scene tag of `scene` (kSlotBits): a new one when the scene has meshes
    // and a tag is free, written to the device table on the stream; 0 otherwise.
    ab_status tagFor(const DevScene& scene, uint32_t& tag);
    void clearTags();
    void freeAll();

    ab_context* ctx_;
    hipStream_t stream_{nullptr};
    void* arena_{nullptr};
    void* accum_{nullptr};   // render accumulators kept between renders (up to kKeepAccumBytes)
    size_t accum_cap_{0};
    uint32_t capacity_{0};
    uint32_t max_jobs_{0};
    PathState ps_{};
    DevJob* d_jobs_{nullptr};
    Counters* h_counters_{nullptr}; // pinned: readCounters' copy, then the two enqueueCounters slots
    hipEvent_t counter_ev_[2]{nullptr, nullptr};
    DevSegment* d_segs_{nullptr};   // segment ring on the device
    DevSegment* h_segs_{nullptr};   // its host mirror (pinned); copies read it when they run
    uint32_t seg_cap_{0};           // ring entries, a power of two
    uint32_t seg_lo_{0}, seg_hi_{0}; // live segments, logical indices
    uint32_t seg_sent_{0};          // segments below this index are on the device
    std::vector<DevJob> jobs_;
    DevScene* d_tags_{nullptr};        // kMaxTags scenes on the device (PathState::tag_scenes)
    std::vector<DevScene> tags_;       // the tags given out, tag t at index t - 1
    unsigned long long work_total_{0};
    LaunchShape shape_{};          // serving: current bounce shape
    SceneSet table_set_{};         // serving: the scenes of the job table
    bool untagged_meshes_{false};  // serving: a job of a scene with meshes got no tag (tags ran out)
    ab_launch_config stream_cfg_{}; // serving: configuration the stream started with
    hipGraphExec_t graph_{nullptr};
    LaunchShape graph_shape_{};     // shape graph_ was captured with
    bool use_graph_{false};
};

// Lock-protected engine of a context for ab_render / ab_render_batch.
Engine& contextEngine(ab_context* ctx);

// The bounce launch shape for a render of one job over `scene` with `capacity`
// path slots under `cfg` (stack class, grid, tracing kernels), as ab_render
// picks it.
LaunchShape singleSceneShape(const ab_launch_config& cfg, uint32_t capacity, const DevScene& scene);

// Most bounces a render of `jobs` (work_total samples over `capacity` slots) can
// take: every path ends within max_depth + 1 bounces of its generation, and a
// bounce generates up to `capacity` paths; running past it means the engine's
// bookkeeping is broken.
uint64_t bounceLimit(const DevJob* jobs, size_t n_jobs, uint64_t work_total, uint32_t capacity);

// Bounce pairs before the first counter check: when paths started in the first
// bounce can have finished, at most 16 bounces in.
uint32_t firstCheckPairs(const DevJob* jobs, size_t n_jobs);

} // namespace ab
