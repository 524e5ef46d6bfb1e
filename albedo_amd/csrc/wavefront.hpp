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
    uint64_t seed;
    uint64_t work_begin; // first global work index of this job
    uint64_t work_count; // width * height * spa^2
    unsigned long long* accum; // 4 per pixel: r, g, b fixed point (2^-28), sample count
    unsigned long long* done;  // samples finished (one counter)
};

// Counters shared by the kernels of one engine.
struct alignas(16) Counters {
    uint32_t n_active;     // paths queued for extension
    uint32_t n_in;         // paths being shaded this bounce
    uint32_t n_shadow;
    uint32_t n_term;
    uint32_t n_free;
    uint32_t n_new;        // paths generate() creates this round
    uint32_t bin_count[8]; // paths per material bin in the last bounce
    uint32_t n_jobs;       // entries of the job table in use
    uint32_t pad;
    unsigned long long work_next;
    unsigned long long work_total;
    unsigned long long camera_rays;
    unsigned long long extend_rays; // every closest-hit ray, camera rays included
    unsigned long long shadow_rays;
    unsigned long long paths_done;
};

constexpr int kBins = 5; // miss, diffuse, conductor, dielectric, mix
constexpr uint32_t kMaxGrid = 8192; // blocks of the grid-stride queue kernels

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
    uint32_t* active;
    uint32_t* binned;
    uint32_t* shadowq;
    uint32_t* termq;
    uint32_t* freeq;
    Counters* ctr;
    const DevJob* jobs;
    uint32_t capacity;
    uint32_t pad;
};

struct LaunchShape {
    uint32_t block;      // 64, 128 or 256
    uint32_t grid;       // blocks of the grid-stride kernels
    uint32_t stack;      // traversal stack depth (max over the jobs' scenes)
};

// Enqueues one bounce: generate, extend, bin, shade, shadow, finish, with the
// bookkeeping kernels between them. No host synchronisation; capturable.
hipError_t launchBounce(const PathState& ps, const LaunchShape& shape, hipStream_t s);
// Initialises counters and the free list for a fresh engine.
hipError_t launchReset(const PathState& ps, unsigned long long work_total, uint32_t n_jobs, hipStream_t s);
// Adds newly admitted work to the total (serving).
hipError_t launchAddWork(const PathState& ps, unsigned long long work_total, uint32_t n_jobs, hipStream_t s);

} // namespace ab
