/*
 * albedo_gpu.h - C ABI of the MI355X path-tracing library (libalbedo_gpu.so).
 *
 * The reference renderer (anazli/albedo) has no C ABI: its seams are C++ types.
 * Each entry point below names the reference interface it stands in for. All
 * functions return an ab_status; none throws. The message for the last failure
 * on a context is available from ab_last_error(). Pointers named d_* are device
 * (HBM) pointers; everything else is host memory owned by the caller.
 *
 * Numerics: traversal and triangle tests run in FP32 with floating-point
 * contraction off. Two steps use FP64: the Woop edge functions are recomputed in
 * FP64 when an FP32 edge function is exactly zero (watertightness), and the sphere
 * quadratic runs in FP64 (radius-1e5 walls lose every digit of c = |o-C|^2 - r^2
 * in FP32). Closest hits are the lexicographic minimum of (t, object, prim), so the
 * result does not depend on BVH layout or launch configuration.
 */
#ifndef ALBEDO_GPU_H
#define ALBEDO_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AB_ABI_VERSION 1

/* Error codes. The reference throws std::invalid_argument, std::out_of_range,
 * std::bad_alloc and SceneParseError (render/render_image.cpp:60-67,
 * scene/scene_file.hpp:40-51); the C++ drop-in shim maps these codes back. */
typedef enum ab_status {
    AB_OK = 0,
    AB_INVALID_ARGUMENT = 1,
    AB_OUT_OF_MEMORY = 2,
    AB_DEVICE_ERROR = 3,
    AB_REJECTED_BUDGET = 4, /* serving: job footprint exceeds the HBM budget */
    AB_NOT_FOUND = 5,       /* unknown job id */
    AB_CANCELLED = 6,       /* job was cancelled */
    AB_PENDING = 7,         /* job not finished yet (poll) */
    AB_PARSE_ERROR = 8      /* scene text rejected (see ab_last_error) */
} ab_status;

typedef struct ab_context ab_context;
typedef struct ab_scene ab_scene;

/* ---------------------------------------------------------------- rays, hits */

/* 32-byte ray: origin, tmin, direction, tmax. Hits are accepted for
 * tmin < t < tmax (reference Interval::surrounds, core/interval.hpp:29). */
typedef struct ab_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} ab_ray;

/* 16-byte hit. object = scene object index in insertion order (reference
 * HitRecord::object, scene/hit_record.hpp:32); prim = triangle index inside the
 * instanced mesh (kept triangles, reference MeshGeometry triangle_list), or
 * AB_NO_PRIM for a sphere. object = AB_MISS when nothing is hit; t is then +inf.
 * bary = barycentrics (b1, b2) of the hit as two unorm16 (b1 in the low half). */
#define AB_MISS 0xffffffffu
#define AB_NO_PRIM 0xffffffffu
typedef struct ab_hit {
    float t;
    uint32_t object;
    uint32_t prim;
    uint32_t bary;
} ab_hit;

/* ------------------------------------------------------------ scene description */

/* A triangle mesh in its own space (reference TriangleMesh,
 * geometry/triangle_mesh.hpp; faces without computable area are dropped and the
 * kept triangles renumbered, triangle_mesh.cpp:76-90). uv: optional corner
 * texture coordinates, 6 floats per face; NULL gives the reference's default
 * (0,0),(1,0),(1,1). smooth != 0 shades with area-weighted vertex normals
 * (triangle_mesh.cpp:92-125, mesh_geometry.cpp:53-64). */
typedef struct ab_mesh_desc {
    const float* positions;   /* 3 * num_vertices */
    uint32_t num_vertices;
    const uint32_t* indices;  /* 3 * num_faces */
    uint32_t num_faces;
    const float* uv;          /* 6 * num_faces, or NULL */
    int32_t smooth;
} ab_mesh_desc;

typedef enum ab_texture_kind {
    AB_TEX_CONSTANT = 0, /* texture/constant_texture.hpp */
    AB_TEX_CHECKER = 1,  /* texture/checker_texture.cpp:8-17 */
    AB_TEX_PERLIN = 2    /* texture/perlin_texture.cpp:7-11 */
} ab_texture_kind;

typedef struct ab_texture_desc {
    uint32_t kind;
    float a[3];      /* constant color / checker even / perlin low */
    float b[3];      /* checker odd / perlin high */
    float frequency; /* checker squares per unit uv / perlin scale */
    float veins;     /* perlin veins (0 = clouds) */
} ab_texture_desc;

typedef enum ab_material_kind {
    AB_MAT_DIFFUSE = 0,    /* material/diffuse_material.hpp:19-23 */
    AB_MAT_CONDUCTOR = 1,  /* material/conductor_material.hpp:27-31 */
    AB_MAT_DIELECTRIC = 2, /* material/dielectric_material.hpp:24-29 */
    AB_MAT_MIX = 3         /* material/mix_material.cpp:7-11 */
} ab_material_kind;

typedef struct ab_material_desc {
    uint32_t kind;
    ab_texture_desc albedo; /* diffuse */
    float eta[3];           /* conductor */
    float k[3];             /* conductor */
    float roughness;        /* conductor, dielectric: alpha = roughness^2 */
    float ior;              /* dielectric */
    uint32_t first, second; /* mix: indices of two non-mix materials */
    float scale, veins, threshold; /* mix: Perlin pattern and threshold */
} ab_material_desc;

typedef enum ab_object_kind { AB_OBJ_SPHERE = 0, AB_OBJ_INSTANCE = 1 } ab_object_kind;

/* One scene object (reference SceneObject, scene/scene_object.hpp:15-22). Spheres
 * keep FP64 center and radius. Instances carry both directions of their affine
 * map, row-major 3x4, as the reference's Transform keeps an exact inverse
 * (core/transform.hpp:38-93). light = index into the light list for an area
 * light's sphere, else -1. */
typedef struct ab_object_desc {
    uint32_t kind;
    uint32_t material;
    int32_t light;
    uint32_t mesh;
    double center[3];
    double radius;
    double to_world[12];
    double to_object[12];
} ab_object_desc;

typedef enum ab_light_kind {
    AB_LIGHT_POINT = 0, /* light/point_light.cpp:7-21 */
    AB_LIGHT_AREA = 1,  /* light/area_light.cpp:31-94: a sphere object that emits */
    AB_LIGHT_SKY = 2    /* light/sky_light.cpp:10-24 */
} ab_light_kind;

typedef struct ab_light_desc {
    uint32_t kind;
    uint32_t object;     /* area: the emitting sphere object */
    double position[3];  /* point */
    float intensity[3];  /* point */
    float radiance[3];   /* area */
    float nadir[3];      /* sky */
    float zenith[3];     /* sky */
} ab_light_desc;

/* The whole scene (reference Scene, scene/scene.hpp:31-98). Lights are listed in
 * the order the reference would have appended them; light choice is uniform over
 * this list (estimator/path_estimator.cpp:54-57, 73-84). */
typedef struct ab_scene_desc {
    const ab_mesh_desc* meshes;
    uint32_t num_meshes;
    const ab_object_desc* objects;
    uint32_t num_objects;
    const ab_material_desc* materials;
    uint32_t num_materials;
    const ab_light_desc* lights;
    uint32_t num_lights;
    uint32_t build_flags;    /* AB_BUILD_* */
} ab_scene_desc;

/* Build the BVH of every mesh with more than 4 triangles on the device (a linear
 * BVH over Morton codes, Karras 2012) instead of the host's binned SAH: tens of
 * milliseconds instead of about a second for a million triangles, for a tree of
 * lower quality (slower traversal, more levels). Query results are the same
 * either way. No counterpart in the reference (accel/bvh.cpp:131-270 builds on
 * the host). */
#define AB_BUILD_DEVICE_BLAS 1u

/* Also build a four-wide form of the host-built trees (each node holds four
 * children: the binary SAH tree collapsed by opening the child of largest
 * surface area) and trace ray queries (ab_trace_closest, ab_trace_any) over it.
 * Results are the same bit for bit. On MI355X it is slower than the binary
 * tree, whose node visit tests two boxes per 64-byte fetch: the queries are
 * bound by instruction issue, a four-wide visit costs more than two binary ones
 * and its stack (up to three entries a visit) takes more LDS. Rejected together
 * with AB_BUILD_DEVICE_BLAS. No counterpart in the reference (accel/bvh.cpp
 * builds a binary tree). */
#define AB_BUILD_WIDE_BVH 2u

/* Trace the scene in one pass. By default the path tracer traces a scene that
 * holds both spheres and mesh instances in two passes: the spheres one ray per
 * lane, then the instances with lane refill, starting from the sphere hit.
 * Images are the same bit for bit either way; the flag is for comparing the
 * two. No counterpart in the reference (scene/scene.cpp:107-133 traces one
 * hierarchy). */
#define AB_BUILD_ONE_PASS 4u

/* Pinhole camera (reference CameraSettings + PerspectiveCamera,
 * camera/perspective_camera.hpp, perspective_camera.cpp:74-110). */
typedef struct ab_camera {
    double look_from[3];
    double look_at[3];
    double up[3];
    double vfov_degrees;
    double aspect;
} ab_camera;

/* Render parameters (reference RenderJob + the RadianceEstimator alternatives,
 * render/render_image.hpp:17-32, estimator/radiance_estimator.hpp:36-38).
 * Each pixel takes samples_per_axis^2 stratified jittered samples.
 *   AB_ESTIMATOR_PATH          PathEstimator{max_depth, roulette_start}
 *                              (estimator/path_estimator.hpp:22-38)
 *   AB_ESTIMATOR_SCATTER_ONLY  ScatterOnlyEstimator{max_scatter_count = max_depth}
 *                              (estimator/scatter_only_estimator.hpp:19-26);
 *                              roulette_start is unused
 *   AB_ESTIMATOR_NORMALS       NormalsEstimator (estimator/normals_estimator.hpp:14-26);
 *                              max_depth and roulette_start are unused */
#define AB_ESTIMATOR_PATH 0u
#define AB_ESTIMATOR_SCATTER_ONLY 1u
#define AB_ESTIMATOR_NORMALS 2u
typedef struct ab_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t samples_per_axis;
    uint32_t max_depth;
    uint32_t roulette_start;
    uint32_t estimator;      /* AB_ESTIMATOR_* */
    uint64_t seed;
} ab_render_params;

/* Per-render counters. */
typedef struct ab_render_stats {
    uint64_t camera_rays;
    uint64_t extension_rays; /* closest-hit rays after the camera ray */
    uint64_t shadow_rays;
    uint64_t paths;
    double gpu_ms;           /* device time of the render, HIP events */
} ab_render_stats;

/* ------------------------------------------------------------------- context */

/* Creates a context on HIP device `device`. */
ab_status ab_context_create(int device, ab_context** out);
void ab_context_destroy(ab_context* ctx);
/* Message of the last failure on this context (thread-local), or "". */
const char* ab_last_error(const ab_context* ctx);
/* Library ABI version (AB_ABI_VERSION). */
int ab_abi_version(void);

/* ------------------------------------------------------------------- scenes */

/* Builds the scene's BVHs on the host and uploads it (reference Scene::add*,
 * scene/scene.cpp:20-105, Bvh::Bvh accel/bvh.cpp:248-270, MeshGeometry
 * mesh_geometry.cpp:68-75). Validation follows the reference constructors
 * (Sphere sphere.cpp:21-41, Triangle triangle.cpp:128-131, MeshInstance
 * mesh_instance.cpp:14-30). A scene with the same content hash as a resident one
 * is shared (reference counted). Limits: at most 29 826 161 kept triangles and
 * 2^26 BVH nodes per scene (AB_INVALID_ARGUMENT beyond). */
ab_status ab_scene_create(ab_context* ctx, const ab_scene_desc* desc, ab_scene** out);
void ab_scene_release(ab_scene* scene);
/* 64-bit content hash of the scene description, and the device bytes it holds. */
uint64_t ab_scene_hash(const ab_scene* scene);
uint64_t ab_scene_device_bytes(const ab_scene* scene);
/* Depth of the deepest BVH and node counts, for tests and reports. */
ab_status ab_scene_info(const ab_scene* scene, uint32_t* max_depth, uint64_t* tlas_nodes,
                        uint64_t* blas_nodes, uint64_t* triangles);

/* -------------------------------------------------------------- ray queries */

/* Closest hit for n rays (reference Scene::intersect, scene/scene.cpp:107-133 ->
 * Bvh::nearestHit accel/bvh.hpp:211-226). d_rays/d_hits are device pointers;
 * the work is enqueued on `stream` (a hipStream_t, NULL = default stream). */
ab_status ab_trace_closest(ab_scene* scene, const ab_ray* d_rays, ab_hit* d_hits, uint64_t n,
                           void* stream);
/* Any hit (reference Scene::occluded, scene/scene.cpp:157-165 -> Bvh::anyHit
 * accel/bvh.hpp:228-237): d_occluded[i] = 1 when ray i hits anything. */
ab_status ab_trace_any(ab_scene* scene, const ab_ray* d_rays, uint8_t* d_occluded, uint64_t n,
                       void* stream);
/* Host-buffer forms of the two queries: copy in, trace, copy out, synchronize. */
ab_status ab_trace_closest_host(ab_scene* scene, const ab_ray* rays, ab_hit* hits, uint64_t n);
ab_status ab_trace_any_host(ab_scene* scene, const ab_ray* rays, uint8_t* occluded, uint64_t n);

/* ------------------------------------------------------------------ rendering */

/* Renders one image with the wavefront path tracer (reference renderImage,
 * render/render_image.cpp:58-77, with PathEstimator::radiance,
 * estimator/path_estimator.cpp:153-185). rgb_mean: host buffer of
 * width*height*3 floats receiving each pixel's mean radiance (reference
 * PixelBuffer::average, image/pixel_buffer.cpp:37-45), row-major from the top-left.
 * stats may be NULL. The image is identical for every launch configuration. */
ab_status ab_render(ab_scene* scene, const ab_camera* camera, const ab_render_params* params,
                    float* rgb_mean, ab_render_stats* stats);

/* Renders one image and ADDS its per-pixel sums to `sums` (host buffer of
 * width*height*4 words in the wire format of ab_job_poll_sums: red, green and
 * blue radiance sums in 2^-AB_SUMS_FRACTION_BITS fixed point, then the sample
 * count), as the reference's renderImage adds samples_per_axis^2 samples to
 * every pixel of the caller's PixelBuffer (render/render_image.hpp:37-48,
 * image/pixel_buffer.cpp:29-35). Rendering again with another seed into the
 * same words accumulates more samples; ab_sums_resolve gives the means, equal
 * to ab_render's for a single render. Words wrap modulo 2^64 only past about
 * 680 000 samples of the largest radiance (1e5) in one pixel. */
ab_status ab_render_sums(ab_scene* scene, const ab_camera* camera, const ab_render_params* params, uint64_t* sums,
                         ab_render_stats* stats);

/* ab_render without allocating: the image goes to d_rgb_mean, a device buffer
 * of width*height*3 floats owned by the caller, on `stream` (a hipStream_t, NULL
 * = default stream), bit-identical to ab_render's. Path state, queues and the
 * per-pixel accumulators live in one array of the library's, 88 MiB on every
 * GPU the library runs on (held until the process ends); the call makes no
 * device, pinned-host or graph allocation. Limits: at most 2^21 pixels (1920 x
 * 1080 fits) -- larger images are rejected with AB_INVALID_ARGUMENT -- and
 * 2^17 paths in flight. Calls that overlap on one GPU (other threads, or
 * streams not ordered after each other) share that array and race. The call
 * returns after the last bounce; the resolve into d_rgb_mean is queued on
 * `stream`. No counterpart in the reference: for callers that cannot allocate
 * (tight serving loops, harnesses that forbid hipMalloc). */
ab_status ab_render_static(ab_scene* scene, const ab_camera* camera, const ab_render_params* params,
                           float* d_rgb_mean, void* stream);

/* Launch-shape knobs of the wavefront engine. Results never depend on them. */
typedef struct ab_launch_config {
    uint32_t block_size;     /* threads per block for the queue kernels: 64, 128 or 256 */
    uint32_t paths_in_flight;/* path slots resident at once */
    uint32_t grid_blocks;    /* blocks per launch of the grid-stride kernels, 0 = auto */
    uint32_t use_graph;      /* capture the per-bounce kernel chain as a hipGraph */
} ab_launch_config;
ab_status ab_set_launch_config(ab_context* ctx, const ab_launch_config* config);
ab_status ab_get_launch_config(const ab_context* ctx, ab_launch_config* config);

/* -------------------------------------------------------- batching / serving */

/* A render job for the batched / served engine. No counterpart in the
 * reference, which renders one RenderJob per call; each job's image equals
 * what ab_render gives for the same scene, camera and params. */
typedef struct ab_job_desc {
    ab_scene* scene;
    ab_camera camera;
    ab_render_params params;
    uint32_t flags;          /* AB_JOB_* (serving only; ignored by ab_render_batch) */
} ab_job_desc;

/* The server keeps a finished job's per-pixel sums (ab_job_poll_sums) until the
 * job is released, in host memory (32 B per pixel). */
#define AB_JOB_KEEP_SUMS 1u

typedef struct ab_job_status {
    uint32_t state;          /* 0 queued, 1 running, 2 done, 3 cancelled, 4 failed */
    uint64_t samples_done;   /* camera samples finished */
    uint64_t samples_total;
    double submit_ms, start_ms, finish_ms; /* host clock, ms since context creation */
} ab_job_status;

/* Renders many jobs together in one persistent wavefront launch per bounce.
 * Images go to rgb_means[i] (width*height*3 floats each). */
ab_status ab_render_batch(ab_context* ctx, const ab_job_desc* jobs, uint32_t n_jobs,
                          float* const* rgb_means, ab_render_stats* stats);

/* Autotuner: times trial renders of `jobs` under candidate launch
 * configurations (block size, paths in flight, grid blocks) and sets the
 * context's configuration to the fastest. Trials keep the jobs' scenes, cameras,
 * widths and samples per pixel and render the same share of every job's rows, so
 * one trial takes about 2^28 camera samples (all of them for smaller job lists).
 * AB_TUNE_BATCH times one ab_render_batch call over all jobs, AB_TUNE_PLAIN one
 * ab_render call per job. `best` and `best_ms` (wall time of
 * the winning trial) may be NULL. Images never depend on the configuration, so
 * tuning changes only speed. No counterpart in the reference. */
#define AB_TUNE_PLAIN 0u
#define AB_TUNE_BATCH 1u
ab_status ab_autotune(ab_context* ctx, const ab_job_desc* jobs, uint32_t n_jobs, uint32_t mode,
                      ab_launch_config* best, double* best_ms);

/* Serving loop: a host thread admits submitted jobs as path slots and the HBM
 * budget allow, renders them together, and publishes progressive results. */
ab_status ab_server_start(ab_context* ctx, uint64_t hbm_budget_bytes);
ab_status ab_server_stop(ab_context* ctx);
/* Order in which the running jobs' samples are started (images never depend on
 * it). The loop keeps about two rounds of bounces' worth of samples handed to the
 * device and refills that window every round:
 *   AB_SCHED_FIFO          in admission order: the least worst-case latency;
 *   AB_SCHED_SHARE         equal shares for every running job (processor sharing):
 *                          small jobs finish fast, large ones stretch;
 *   AB_SCHED_SMALL_FIRST   the jobs with the fewest samples left first (shortest
 *                          remaining work first): a low median latency with few
 *                          jobs sharing the device at a time (the default).
 * Takes effect at the next round. No counterpart in the reference. */
#define AB_SCHED_FIFO 0u
#define AB_SCHED_SHARE 1u
#define AB_SCHED_SMALL_FIRST 2u
ab_status ab_server_set_policy(ab_context* ctx, uint32_t policy);
ab_status ab_job_submit(ab_context* ctx, const ab_job_desc* job, uint64_t* job_id);
/* HBM bytes a job holds while admitted (accumulators, counter, mean image): the
 * amount ab_job_submit charges against the budget. Validates the job like
 * ab_job_submit. No counterpart in the reference; for callers that plan budgets. */
ab_status ab_job_footprint(const ab_job_desc* job, uint64_t* bytes);
/* Progressive snapshot: the mean over the samples finished so far. */
ab_status ab_job_poll(ab_context* ctx, uint64_t job_id, ab_job_status* status, float* rgb_mean);

/* Progressive results as exact sums (the wire format). Per pixel, row-major from
 * the top-left corner, four little-endian 64-bit words: the red, green and blue
 * radiance sums as unsigned fixed point (value = word / 2^AB_SUMS_FRACTION_BITS)
 * and the sample count -- the reference's per-pixel accumulator of sum and count
 * (image/pixel_buffer.hpp:54-57) in integer form. Sums of disjoint sample sets
 * combine exactly, in any order, by adding words (ab_sums_merge), and
 * ab_sums_resolve gives the same means as ab_job_poll / ab_job_wait. A snapshot
 * of a running job may catch a pixel between its sum and its count update; the
 * final one (after ab_job_wait) is exact. A finished job answers only if it was
 * submitted with AB_JOB_KEEP_SUMS. No counterpart in the reference. */
#define AB_SUMS_FRACTION_BITS 28
ab_status ab_job_poll_sums(ab_context* ctx, uint64_t job_id, ab_job_status* status, uint64_t* sums);
ab_status ab_sums_merge(uint64_t* dst, const uint64_t* src, uint64_t pixels);
ab_status ab_sums_resolve(const uint64_t* sums, uint64_t pixels, float* rgb_mean);
ab_status ab_job_wait(ab_context* ctx, uint64_t job_id, ab_job_status* status, float* rgb_mean);
ab_status ab_job_cancel(ab_context* ctx, uint64_t job_id);
ab_status ab_job_release(ab_context* ctx, uint64_t job_id);

typedef struct ab_server_stats {
    uint64_t jobs_done;
    uint64_t jobs_rejected;
    uint64_t rays;           /* camera + extension + shadow rays traced */
    double busy_ms;
    double p50_latency_ms;
    double p99_latency_ms;
    uint64_t hbm_charged;    /* bytes of the admitted jobs' device buffers, held against the budget */
} ab_server_stats;
ab_status ab_server_stats_get(ab_context* ctx, ab_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* ALBEDO_GPU_H */
