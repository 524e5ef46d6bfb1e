// Wavefront path-tracing engine: device-side job table, path state and queues,
// and the host-callable launchers of its kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include "layout.hpp"

namespace ab {

// One render job resident on the device.
struct alignas(16) DevJob {
    DevScene scene;
    float cam_origin[4];
    float cam_forward[4];
    float cam_hw[4];
    float cam_hh[4];
    uint32_t width, height, spa, max_depth;
    uint32_t roulette_start;
    uint32_t cancelled;
    uint32_t group;     // samples of one pixel that are consecutive in work order
    uint32_t estimator; // AB_ESTIMATOR_*
    uint32_t tag;       // scene tag of the job's queue entries (see kSlotBits), 0 for none
    uint64_t seed;
    uint64_t work_count; // width * height * spa^2
    unsigned long long* accum; // 4 per pixel: r, g, b fixed point (2^-28), sample count
    unsigned long long* done;  // samples finished: the sum of kDoneLanes counters (see below)
};

// A job's finished-sample count is spread over kDoneLanes counters kDoneStride
// entries (256 B) apart, block b adding to counter b mod kDoneLanes: atomics on
// one address serialise at about 12 ns each on MI355X, which made a single
// counter cost ~100 us per half a million finished paths.
constexpr uint32_t kDoneLanes = 64;
constexpr uint32_t kDoneStride = 32;
constexpr size_t kDoneBytes = size_t(kDoneLanes) * kDoneStride * 8;

// A run of one job's work placed at global work indices [begin, next segment's
// begin): global index w is the job's work index local + (w - begin). Generation
// walks global indices in order, so the order of segments is the order in which
// jobs' samples start; the serving loop interleaves jobs by appending short
// segments of each. Segments live in a ring of a power-of-two size, live range
// [seg_lo, seg_hi) in logical (unwrapped) indices. Images do not depend on how a
// job's work is cut, since a sample's random numbers depend only on its job,
// pixel and index.
struct alignas(8) DevSegment {
    uint64_t begin;
    uint64_t local;
    uint32_t job;
    uint32_t pad;
};

// Counters shared by the kernels of one engine.
// n_term and n_shadow of a bounce are added to paths_done and shadow_rays (and
// the terminated paths' slots that are not recycled to n_free) at the next
// bounce's kPlanBins; the host adds what is still pending when it reads the
// block (readCounters).
//
// Recycling: while the work left covers every path of a bounce, a terminated
// path keeps its place in the next active queue and the next bounce's first new
// samples take over its slot and queue position (n_recycle of them, in
// termination-queue order). The active queue then keeps one order for the whole
// render -- ascending slots from the first generation on -- so neighbouring
// lanes read neighbouring slots, instead of new paths taking freed slots that
// lie anywhere in the state arrays.
struct alignas(16) Counters {
    uint32_t n_active;     // paths queued for extension (with recycling: the kept positions too)
    uint32_t n_in;         // paths being shaded this bounce
    uint32_t n_shadow;
    uint32_t n_term;
    uint32_t n_free;
    uint32_t n_recycle;    // terminated paths whose slot and position the next new samples take
    uint32_t bin_count[8]; // paths per material bin in the last bounce
    uint32_t seg_lo;       // live segments, logical ring indices [seg_lo, seg_hi)
    uint32_t seg_hi;
    unsigned long long work_next;
    unsigned long long work_total;
    unsigned long long camera_rays;
    unsigned long long extend_rays; // every closest-hit ray, camera rays included
    unsigned long long shadow_rays;
    unsigned long long paths_done;
    // Split tracing: the extension queue's positions are sorted by scene tag
    // (`bytag`), the untagged ones first; the shadow queue comes out in that
    // order too. Untagged entries of each: [0, n_untagged), [0, sh_untagged).
    uint32_t n_untagged;
    uint32_t sh_untagged;
    // The free stack pops slots in ascending order (set by a reset and by the
    // free-slot sort, cleared when a finish pushes slots).
    uint32_t free_sorted;
    uint32_t pad1;
};

// Queue entries (active, shadow queues) hold a path's slot in their low
// kSlotBits bits and, above them, the scene tag of its job: 0 for paths whose
// rays go one per lane (kExtend, kShadow), or t in 1..kMaxTags-1 for paths of the
// scene tag_scenes[t], whose rays kTraceGrouped traces a scene at a time. Copying
// an entry from queue to queue keeps the tag; the free list holds bare slots and
// the termination queue positions in `active`. Paths in flight are at most 2^kSlotBits.
constexpr uint32_t kSlotBits = 26;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
constexpr uint32_t kMaxTags = 1u << (32 - kSlotBits);

constexpr int kBins = 5; // miss, diffuse, conductor, dielectric, mix
constexpr uint32_t kMaxGrid = 8192; // blocks of the grid-stride queue kernels
// The tag sort's kernels: kTagGrid blocks of kTagBlock threads, each block a
// contiguous range of the queue.
constexpr uint32_t kTagGrid = 512;
constexpr uint32_t kTagBlock = 256;
// The free-slot sort's kernels: kFreeGrid blocks of 256 threads, each a
// contiguous range of the slots.
constexpr uint32_t kFreeGrid = 256;

// Path state, structure of arrays over `capacity` slots.
struct PathState {
    float4* org;    // xyz origin of the next ray, w arrival pdf
    float4* dir;    // xyz direction, w = 1 when the arrival was specular
    float4* thr;    // xyz throughput, w depth (int bits)
    float4* rad;    // xyz radiance gathered so far, w unused
    float4* prev;   // xyz arrival origin, w job index (int bits)
    uint64_t* rng;  // PCG32 state (the increment is recomputed from the pixel)
    uint32_t* pixel;
    uint4* hit;     // t bits, object, prim, b1 bits
    float* hit_b2;
    uint32_t* key;  // per queue position: the output queues of the item (bit mask)
    uint32_t* block_count; // per block of a queue kernel: items per output queue (8 per block)
    uint32_t* block_off;   // per block: first output position per queue
    float4* sh_org; // xyz origin, w tmax
    float4* sh_dir; // xyz direction
    float4* sh_c;   // xyz contribution if unoccluded
    uint32_t* active;      // paths to extend (queue entries; see Counters on their order)
    uint32_t* active_next; // the next bounce's active queue (the two swap every bounce)
    uint32_t* binned;      // positions in `active`, grouped by the material hit
    uint32_t* shadowq;
    uint32_t* termq;       // positions in `active` of the terminated paths
    uint32_t* freeq;       // free slots (a stack)
    uint8_t* slot_free;    // per slot: 1 while it is on the free stack
    uint32_t* bytag;       // split tracing: positions in `active`, sorted by scene tag
    uint32_t* tag_count;   // split tracing: per tag and tag-sort block, entries, then first output index
    Counters* ctr;
    const DevJob* jobs;
    const DevScene* tag_scenes; // kMaxTags entries: the scene of each tag (entry 0 unused)
    const DevSegment* segs; // ring of seg_mask + 1 entries
    uint32_t capacity;
    uint32_t seg_mask;
};

// How the extension and shadow rays are traced: one ray per lane (kernels
// kExtend, kShadow); with lane refill (kTraceQueue) over the one scene of every
// job; or split, the rays of scenes with meshes with lane refill grouped by scene
// (kTraceGrouped: a wave holds one scene at a time) and the others one per lane.
enum TraceMode : uint32_t { kTracePlain = 0, kTraceOne = 1, kTraceSplit = 2 };

struct LaunchShape {
    uint32_t block;        // 64, 128 or 256
    uint32_t grid;         // blocks of the grid-stride kernels
    uint32_t stack;        // traversal stack class (max over the jobs' scenes)
    uint32_t pstack{0};    // stack class of the one-ray-per-lane kernels (max over the scenes they trace)
    uint32_t refill{kTracePlain};
    uint32_t qgrid{0};     // one-wave blocks of the lane-refill kernels
    uint32_t qstack{0};    // stack class of the lane-refill kernels (20, 24, 32 or 64 entries)
    uint32_t q_spheres{0}; // a scene the refill kernels trace has spheres (the FP64 sphere test is compiled in)
    uint32_t q_reg{0};     // such a scene's stack bound + 1 exceeds the class: TLAS leaf rest in a register
    DevScene scene{};      // kTraceOne: the scene (its instance-only TLAS with sphere_pass)
    // kTraceOne over a scene with both spheres and instances: kSpherePass traces
    // `sphere_scene` (its sphere-only TLAS, stack class `sstack`) before the
    // refill kernel traces `scene` seeded with the sphere hits.
    uint32_t sphere_pass{0};
    uint32_t sstack{0};
    DevScene sphere_scene{};
    // 0 when no bounce can append new samples after the first (all of a render's
    // samples fit the paths in flight): the free-slot sort is left out, as it
    // would find nothing to do.
    uint32_t free_sort{1};
    // 0 for the bounces of a render after its first pair when every sample
    // started in the first bounce: the generation kernel is left out.
    uint32_t generate{1};
};

// Enqueues two bounces (generate, extend, bin, shade, shadow, finish, with the
// bookkeeping kernels between them): the first reads `active` and writes
// `active_next`, the second the other way round, so the queues are back in
// place afterwards. No host synchronisation; capturable.
hipError_t launchBouncePair(const PathState& ps, const LaunchShape& shape, hipStream_t s);
// Initialises counters and the free list for a fresh engine whose work is
// `work_total` samples in segments [0, n_segs) of the ring.
hipError_t launchReset(const PathState& ps, unsigned long long work_total, uint32_t n_segs, hipStream_t s);
// Publishes segments appended to the ring: work now ends at `work_total`, live
// segments are [seg_lo, seg_hi) (serving).
hipError_t launchAddWork(const PathState& ps, unsigned long long work_total, uint32_t seg_lo, uint32_t seg_hi,
                         hipStream_t s);
// Per-pixel mean radiance of a job: sums divided by the sample count, into
// mean[npx * 3] (zero where no sample landed yet).
hipError_t launchResolve(const unsigned long long* accum, uint64_t npx, float* mean, hipStream_t s);
// Finished-sample counts of up to kDoneListMax jobs: out[i] = the sum of the
// kDoneLanes counters at done[i] (a job's `done` lanes), one wave per job.
constexpr uint32_t kDoneListMax = 64;
hipError_t launchSumDone(unsigned long long* const* done, uint32_t n, unsigned long long* out, hipStream_t s);

} // namespace ab
