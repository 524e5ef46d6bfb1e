// C ABI: contexts, scenes and ray queries.

#include "blas_build.hpp"
#include "context.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

namespace ab {

namespace {
thread_local std::string t_last_error;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
} // namespace

ab_status fail(ab_status code, const std::string& msg)
{
    t_last_error = msg;
    return code;
}

ab_status hipFail(hipError_t e, const char* what)
{
    const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return fail(e == hipErrorOutOfMemory ? AB_OUT_OF_MEMORY : AB_DEVICE_ERROR, msg);
}

double msSince(const ab_context* ctx)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ctx->t0).count();
}

} // namespace ab

using namespace ab;

extern "C" {

int ab_abi_version(void) { return AB_ABI_VERSION; }

const char* ab_last_error(const ab_context*) { return t_last_error.c_str(); }

ab_status ab_context_create(int device, ab_context** out)
{
    if (!out) return fail(AB_INVALID_ARGUMENT, "ab_context_create: out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hipFail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count)
        return fail(AB_INVALID_ARGUMENT, "ab_context_create: device " + std::to_string(device) +
                                             " out of range (" + std::to_string(count) + " devices)");
    if ((e = hipSetDevice(device)) != hipSuccess) return hipFail(e, "hipSetDevice");
    staticScratchSetup(device);
    ab_context* ctx = new (std::nothrow) ab_context;
    if (!ctx) return fail(AB_OUT_OF_MEMORY, "ab_context_create: out of host memory");
    ctx->device = device;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete ctx;
        return hipFail(e, "hipStreamCreate");
    }
    *out = ctx;
    return AB_OK;
}

void ab_context_destroy(ab_context* ctx) { delete ctx; }

ab_status ab_scene_create(ab_context* ctx, const ab_scene_desc* desc, ab_scene** out)
{
    if (!ctx || !desc || !out) return fail(AB_INVALID_ARGUMENT, "ab_scene_create: null argument");
    *out = nullptr;
    const uint64_t hash = hashSceneDesc(*desc);
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->resident.find(hash);
        if (it != ctx->resident.end()) {
            it->second->refs.fetch_add(1);
            *out = it->second;
            return AB_OK;
        }
    }
    HostScene hs;
    std::string err;
    try {
        const ab_status st = buildHostScene(*desc, hs, err);
        if (st != AB_OK) return fail(st, err);
    } catch (const std::bad_alloc&) {
        return fail(AB_OUT_OF_MEMORY, "ab_scene_create: out of host memory building the BVH");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    // Device-built BLASes (AB_BUILD_DEVICE_BLAS): sort and link on the device,
    // then place their nodes after the host-built ones.
    std::vector<std::unique_ptr<DeviceBlas>> dblas;
    size_t device_nodes = 0;
    for (const DeviceMesh& dm : hs.device_meshes) {
        auto b = std::make_unique<DeviceBlas>();
        if ((e = b->prepare(hs.tris.data() + dm.tri_base, dm.count, ctx->stream)) != hipSuccess)
            return hipFail(e, "ab_scene_create: device BVH build");
        const uint32_t root = uint32_t(hs.nodes.size() + device_nodes);
        for (DevObject& o : hs.objects)
            if (o.kind == kObjInstance && o.mesh == dm.mesh) o.blas_root = int32_t(root);
        hs.max_blas_depth = std::max(hs.max_blas_depth, b->levels());
        hs.blas_nodes += b->nodes();
        device_nodes += b->nodes();
        dblas.push_back(std::move(b));
    }
    if (hs.tris.size() > size_t(kMaxTriangles))
        return fail(AB_INVALID_ARGUMENT, "ab_scene_create: too many triangles (at most " +
                                             std::to_string(kMaxTriangles) + ")");
    if (hs.nodes.size() + device_nodes >= size_t(kMaxNodes))
        return fail(AB_INVALID_ARGUMENT, "ab_scene_create: too many BVH nodes (at most " +
                                             std::to_string(kMaxNodes) + ")");
    // A traversal pushes one entry per interior node on its current root-to-leaf
    // path, so it needs at most (levels - 1) entries per tree; a BLAS traversal
    // starts on top of the TLAS entries pending at the instance's leaf (the rest of
    // that leaf waits in a register, not on the stack).
    const uint32_t stack =
        (hs.tlas_depth ? hs.tlas_depth - 1 : 0) + (hs.max_blas_depth ? hs.max_blas_depth - 1 : 0);
    if (stack > kMaxStackDepth)
        return fail(AB_INVALID_ARGUMENT, "ab_scene_create: BVH too deep for the traversal stack");
    // The four-wide form (AB_BUILD_WIDE_BVH) pushes up to three entries a visit.
    const uint32_t wide_stack = hs.wide_tlas_stack + hs.wide_blas_stack;
    if (hs.wide_tlas_root != kNoChild && wide_stack > kMaxStackDepth)
        return fail(AB_INVALID_ARGUMENT, "ab_scene_create: four-wide BVH too deep for the traversal stack");

    // One arena holds every array, each 256-byte aligned. Triangles take three
    // copies (DevScene::tris): the built ones go into the third, whose
    // coordinate order is the built (x, y, z), and a kernel rotates the others.
    const size_t n_tris = hs.tris.size();
    struct Part {
        const void* src;
        size_t bytes; // copied from src
        size_t off;
        size_t room;  // bytes the part occupies (>= at + bytes)
        size_t at;    // where in the part the copy goes
    };
    Part parts[] = {{hs.nodes.data(), hs.nodes.size() * sizeof(DevNode), 0,
                     (hs.nodes.size() + device_nodes) * sizeof(DevNode), 0},
                    {hs.tris.data(), n_tris * sizeof(DevTri), 0, 3 * n_tris * sizeof(DevTri),
                     2 * n_tris * sizeof(DevTri)},
                    {hs.leafobj.data(), hs.leafobj.size() * sizeof(uint32_t), 0, 0, 0},
                    {hs.objects.data(), hs.objects.size() * sizeof(DevObject), 0, 0, 0},
                    {hs.spheres.data(), hs.spheres.size() * sizeof(double), 0, 0, 0},
                    {hs.shading.data(), hs.shading.size() * sizeof(DevTriShading), 0, 0, 0},
                    {hs.materials.data(), hs.materials.size() * sizeof(DevMaterial), 0, 0, 0},
                    {hs.lights.data(), hs.lights.size() * sizeof(DevLight), 0, 0, 0}};
    size_t total = 0;
    for (Part& p : parts) {
        p.room = std::max(p.room, p.at + p.bytes);
        p.off = total;
        total += align256(std::max<size_t>(p.room, 16));
    }
    void* arena = nullptr;
    if ((e = hipMalloc(&arena, total)) != hipSuccess) return hipFail(e, "hipMalloc(scene)");
    // Each array straight from its host vector (no staging copy); padding between
    // parts and the device-built nodes' room are never read before written.
    for (const Part& p : parts) {
        if (!p.bytes) continue;
        if ((e = hipMemcpy(static_cast<unsigned char*>(arena) + p.off + p.at, p.src, p.bytes,
                           hipMemcpyHostToDevice)) != hipSuccess) {
            (void)hipFree(arena);
            return hipFail(e, "hipMemcpy(scene)");
        }
    }
    {
        // Device-built BLASes: nodes after the host's, triangles over their
        // kept-order copies, in leaf order; then the two rotated copies.
        auto* nodes = reinterpret_cast<DevNode*>(static_cast<unsigned char*>(arena) + parts[0].off);
        auto* tris = reinterpret_cast<DevTri*>(static_cast<unsigned char*>(arena) + parts[1].off);
        DevTri* built = tris + 2 * n_tris;
        size_t base = hs.nodes.size();
        for (size_t i = 0; i < dblas.size() && e == hipSuccess; ++i) {
            const DeviceMesh& dm = hs.device_meshes[i];
            e = dblas[i]->emit(nodes + base, uint32_t(base), built + dm.tri_base, dm.tri_base, ctx->stream);
            base += dblas[i]->nodes();
        }
        if (e == hipSuccess) e = launchRotateTris(tris, uint32_t(n_tris), ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        dblas.clear();
        if (e != hipSuccess) {
            (void)hipFree(arena);
            return hipFail(e, "ab_scene_create: device BVH emission");
        }
    }
    ab_scene* s = new (std::nothrow) ab_scene;
    if (!s) {
        (void)hipFree(arena);
        return fail(AB_OUT_OF_MEMORY, "ab_scene_create: out of host memory");
    }
    unsigned char* base = static_cast<unsigned char*>(arena);
    s->ctx = ctx;
    s->hash = hash;
    s->arena = arena;
    s->bytes = total;
    s->dev.nodes = reinterpret_cast<const DevNode*>(base + parts[0].off);
    s->dev.tris = reinterpret_cast<const DevTri*>(base + parts[1].off);
    s->dev.leafobj = reinterpret_cast<const uint32_t*>(base + parts[2].off);
    s->dev.objects = reinterpret_cast<const DevObject*>(base + parts[3].off);
    s->dev.spheres = reinterpret_cast<const double*>(base + parts[4].off);
    s->dev.shading = reinterpret_cast<const DevTriShading*>(base + parts[5].off);
    s->dev.materials = reinterpret_cast<const DevMaterial*>(base + parts[6].off);
    s->dev.lights = reinterpret_cast<const DevLight*>(base + parts[7].off);
    s->dev.tlas_root = hs.tlas_root;
    s->dev.num_objects = uint32_t(hs.objects.size());
    s->dev.num_lights = uint32_t(hs.lights.size());
    s->dev.sky_light = hs.sky_light;
    s->dev.stack_depth = stack;
    s->dev.tri_count = uint32_t(n_tris);
    // The four-wide form covers host-built trees only.
    s->dev.wide_root = hs.wide_tlas_root;
    s->dev.wide_stack = wide_stack;
    // Kind-split TLASes (spheres one ray per lane, then instances seeded with the
    // sphere hit): stack bounds as for the whole TLAS; the path tracer's stack
    // classes reach 64 entries, beyond which the scene is traced whole.
    s->dev.sphere_stack = hs.sphere_depth ? hs.sphere_depth - 1 : 0;
    s->dev.mesh_stack = (hs.mesh_depth ? hs.mesh_depth - 1 : 0) + (hs.max_blas_depth ? hs.max_blas_depth - 1 : 0);
    const bool kind_split = hs.sphere_root != kNoChild && hs.mesh_root != kNoChild && s->dev.sphere_stack <= 64 &&
                            s->dev.mesh_stack <= 64;
    s->dev.sphere_root = kind_split ? hs.sphere_root : kNoChild;
    s->dev.mesh_root = kind_split ? hs.mesh_root : kNoChild;
    s->dev.has_spheres = 0;
    s->dev.has_meshes = 0;
    for (const DevObject& o : hs.objects) {
        s->dev.has_spheres |= o.kind == kObjSphere ? 1u : 0u;
        s->dev.has_meshes |= o.kind == kObjSphere ? 0u : 1u;
    }
    s->tlas_depth = hs.tlas_depth;
    s->blas_depth = hs.max_blas_depth;
    s->tlas_nodes = hs.tlas_nodes;
    s->blas_nodes = hs.blas_nodes;
    s->triangles = hs.tris.size();
    s->num_lights = uint32_t(hs.lights.size());
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->resident.find(hash);
        if (it != ctx->resident.end()) {
            // Another thread made the same scene meanwhile: share theirs.
            it->second->refs.fetch_add(1);
            *out = it->second;
            (void)hipFree(arena);
            delete s;
            return AB_OK;
        }
        ctx->resident[hash] = s;
    }
    *out = s;
    return AB_OK;
}

void ab_scene_release(ab_scene* s)
{
    if (!s) return;
    ab_context* ctx = s->ctx;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        if (s->refs.fetch_sub(1) != 1) return;
        ctx->resident.erase(s->hash);
    }
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(s->arena);
    delete s;
}

uint64_t ab_scene_hash(const ab_scene* s) { return s ? s->hash : 0; }
uint64_t ab_scene_device_bytes(const ab_scene* s) { return s ? s->bytes : 0; }

ab_status ab_scene_info(const ab_scene* s, uint32_t* max_depth, uint64_t* tlas_nodes, uint64_t* blas_nodes,
                        uint64_t* triangles)
{
    if (!s) return fail(AB_INVALID_ARGUMENT, "ab_scene_info: null scene");
    if (max_depth) *max_depth = s->dev.stack_depth;
    if (tlas_nodes) *tlas_nodes = s->tlas_nodes;
    if (blas_nodes) *blas_nodes = s->blas_nodes;
    if (triangles) *triangles = s->triangles;
    return AB_OK;
}

ab_status ab_trace_closest(ab_scene* s, const ab_ray* d_rays, ab_hit* d_hits, uint64_t n, void* stream)
{
    if (!s) return fail(AB_INVALID_ARGUMENT, "ab_trace_closest: null scene");
    if (n && (!d_rays || !d_hits)) return fail(AB_INVALID_ARGUMENT, "ab_trace_closest: null buffer");
    if ((reinterpret_cast<uintptr_t>(d_rays) | reinterpret_cast<uintptr_t>(d_hits)) & 15u)
        return fail(AB_INVALID_ARGUMENT, "ab_trace_closest: buffers must be 16-byte aligned");
    const hipError_t e = launchTraceClosest(s->dev, d_rays, d_hits, n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AB_OK : hipFail(e, "trace_closest launch");
}

ab_status ab_trace_any(ab_scene* s, const ab_ray* d_rays, uint8_t* d_occ, uint64_t n, void* stream)
{
    if (!s) return fail(AB_INVALID_ARGUMENT, "ab_trace_any: null scene");
    if (n && (!d_rays || !d_occ)) return fail(AB_INVALID_ARGUMENT, "ab_trace_any: null buffer");
    if (reinterpret_cast<uintptr_t>(d_rays) & 15u)
        return fail(AB_INVALID_ARGUMENT, "ab_trace_any: rays must be 16-byte aligned");
    const hipError_t e = launchTraceAny(s->dev, d_rays, d_occ, n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AB_OK : hipFail(e, "trace_any launch");
}

} // extern "C"

namespace {
template <class Out, class Launch>
ab_status traceHost(ab_scene* s, const ab_ray* rays, Out* out, uint64_t n, Launch launch, const char* what)
{
    if (!s) return fail(AB_INVALID_ARGUMENT, std::string(what) + ": null scene");
    if (n == 0) return AB_OK;
    if (!rays || !out) return fail(AB_INVALID_ARGUMENT, std::string(what) + ": null buffer");
    ab_context* ctx = s->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    void *dr = nullptr, *dout = nullptr;
    if ((e = hipMallocAsync(&dr, n * sizeof(ab_ray), ctx->stream)) != hipSuccess) return hipFail(e, what);
    if ((e = hipMallocAsync(&dout, n * sizeof(Out), ctx->stream)) != hipSuccess) {
        (void)hipFreeAsync(dr, ctx->stream);
        return hipFail(e, what);
    }
    e = hipMemcpyAsync(dr, rays, n * sizeof(ab_ray), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch(s->dev, static_cast<const ab_ray*>(dr), static_cast<Out*>(dout), n, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * sizeof(Out), hipMemcpyDeviceToHost, ctx->stream);
    (void)hipFreeAsync(dr, ctx->stream);
    (void)hipFreeAsync(dout, ctx->stream);
    const hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hipFail(e, what);
    if (e2 != hipSuccess) return hipFail(e2, what);
    return AB_OK;
}
} // namespace

extern "C" {

ab_status ab_trace_closest_host(ab_scene* s, const ab_ray* rays, ab_hit* hits, uint64_t n)
{
    return traceHost(s, rays, hits, n, launchTraceClosest, "ab_trace_closest_host");
}

ab_status ab_trace_any_host(ab_scene* s, const ab_ray* rays, uint8_t* occ, uint64_t n)
{
    return traceHost(s, rays, occ, n, launchTraceAny, "ab_trace_any_host");
}

} // extern "C"
