"""albedo_amd: MI355X path-tracing hot path behind a C ABI (libalbedo_gpu.so).

Python mirror of the reference renderer's interface for tests and benchmarks:
  Context         device context (ab_context_create)
  GpuScene        a resident scene (reference Scene after its add* calls)
  render_image    reference renderImage(RenderJob, PixelBuffer)
                  (render/render_image.cpp:58-77)
All compute goes through the HIP library; there is no CPU fallback.
"""
import ctypes
import weakref

import numpy as np

from . import _abi
from ._abi import AlbedoError, check, lib
from .scene import (CameraSettings, CheckerTexture, ConductorMaterial, ConstantTexture, DielectricMaterial,
                    DiffuseMaterial, MeshGeometry, MeshInstance, MixMaterial, PerlinPattern, PerlinTexture,
                    PointLight, Scene, SkyLight, Sphere, Transform, TriangleMesh, FLAT, SMOOTH)

__all__ = ["Context", "GpuScene", "AlbedoError", "Scene", "Sphere", "TriangleMesh", "MeshGeometry",
           "MeshInstance", "Transform", "DiffuseMaterial", "ConductorMaterial", "DielectricMaterial",
           "MixMaterial", "PerlinPattern", "ConstantTexture", "CheckerTexture", "PerlinTexture", "PointLight",
           "SkyLight", "CameraSettings", "FLAT", "SMOOTH"]


class Context:
    """A library context on one HIP device."""

    def __init__(self, device=0):
        self.lib = lib()
        h = ctypes.c_void_p()
        check(self.lib.ab_context_create(device, ctypes.byref(h)))
        self.handle = h
        self._scenes = weakref.WeakSet()

    def close(self):
        if self.handle:
            # Destroying the context frees its scenes: their handles die with it.
            for s in list(self._scenes):
                s.handle = None
            self.lib.ab_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status):
        check(status, self.handle)

    def scene(self, scene: Scene, device_blas=False, wide_bvh=False, one_pass=False) -> "GpuScene":
        """Uploads `scene`; device_blas builds its mesh BVHs on the GPU (AB_BUILD_DEVICE_BLAS),
        wide_bvh traces ray queries over four-wide trees (AB_BUILD_WIDE_BVH), one_pass
        has the path tracer trace every object in one pass (AB_BUILD_ONE_PASS)."""
        return GpuScene(self, scene, device_blas, wide_bvh, one_pass)

    def set_launch_config(self, block_size=256, paths_in_flight=1 << 21, grid_blocks=0, use_graph=1):
        c = _abi.LaunchConfig(block_size, paths_in_flight, grid_blocks, use_graph)
        self.check(self.lib.ab_set_launch_config(self.handle, ctypes.byref(c)))

    def launch_config(self):
        c = _abi.LaunchConfig()
        self.check(self.lib.ab_get_launch_config(self.handle, ctypes.byref(c)))
        return c

    def autotune(self, jobs, batched=True):
        """Picks and sets the fastest launch configuration for `jobs` (tuples as
        for render_batch). Returns (config dict, trial wall ms)."""
        descs = (_abi.JobDesc * len(jobs))()
        for i, j in enumerate(jobs):
            descs[i] = Server.job_desc(*j)
        c = _abi.LaunchConfig()
        ms = ctypes.c_double()
        self.check(self.lib.ab_autotune(self.handle, descs, len(jobs),
                                        _abi.AB_TUNE_BATCH if batched else _abi.AB_TUNE_PLAIN, ctypes.byref(c),
                                        ctypes.byref(ms)))
        return {"block_size": c.block_size, "paths_in_flight": c.paths_in_flight, "grid_blocks": c.grid_blocks,
                "use_graph": c.use_graph}, ms.value


class GpuScene:
    """A scene resident in HBM (ab_scene_create)."""

    def __init__(self, ctx: Context, scene: Scene, device_blas=False, wide_bvh=False, one_pass=False):
        self.ctx = ctx
        self.desc = scene.to_desc()
        self.desc.desc.build_flags = ((_abi.BUILD_DEVICE_BLAS if device_blas else 0) |
                                      (_abi.BUILD_WIDE_BVH if wide_bvh else 0) |
                                      (_abi.BUILD_ONE_PASS if one_pass else 0))
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.ab_scene_create(ctx.handle, self.desc.ptr(), ctypes.byref(h)))
        self.handle = h
        ctx._scenes.add(self)

    def release(self):
        if self.handle and self.ctx.handle:
            self.ctx.lib.ab_scene_release(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    @property
    def hash(self):
        return self.ctx.lib.ab_scene_hash(self.handle)

    @property
    def device_bytes(self):
        return self.ctx.lib.ab_scene_device_bytes(self.handle)

    def info(self):
        d = ctypes.c_uint32()
        t, b, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.ab_scene_info(self.handle, ctypes.byref(d), ctypes.byref(t), ctypes.byref(b),
                                                  ctypes.byref(n)))
        return {"stack_depth": d.value, "tlas_nodes": t.value, "blas_nodes": b.value, "triangles": n.value}

    # ------------------------------------------------------------ ray queries

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        """Host rays [n, 8] float32 -> hits [n, 4] uint32 (t bits, object, prim, bary)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.zeros((len(rays), 4), dtype=np.uint32)
        self.ctx.check(self.ctx.lib.ab_trace_closest_host(self.handle, rays.ctypes.data, hits.ctypes.data,
                                                          len(rays)))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(len(rays), dtype=np.uint8)
        self.ctx.check(self.ctx.lib.ab_trace_any_host(self.handle, rays.ctypes.data, occ.ctypes.data, len(rays)))
        return occ

    def trace_closest_device(self, d_rays: int, d_hits: int, n: int, stream: int = 0):
        """Device pointers (e.g. torch tensor data_ptr()) on a hipStream_t."""
        self.ctx.check(self.ctx.lib.ab_trace_closest(self.handle, ctypes.c_void_p(d_rays), ctypes.c_void_p(d_hits),
                                                     n, ctypes.c_void_p(stream)))

    def trace_any_device(self, d_rays: int, d_occ: int, n: int, stream: int = 0):
        self.ctx.check(self.ctx.lib.ab_trace_any(self.handle, ctypes.c_void_p(d_rays), ctypes.c_void_p(d_occ), n,
                                                 ctypes.c_void_p(stream)))

    # -------------------------------------------------------------- rendering

    def render(self, camera: CameraSettings, width, height, samples_per_axis, max_depth=64, roulette_start=3,
               seed=0, stats=False, estimator=_abi.ESTIMATOR_PATH):
        """Mean radiance per pixel, float32 [height, width, 3]. estimator: one of
        _abi.ESTIMATOR_PATH / ESTIMATOR_SCATTER_ONLY (max_depth = scatter limit) /
        ESTIMATOR_NORMALS."""
        from .workloads import render_params
        p = render_params(width, height, samples_per_axis, max_depth, roulette_start, seed, estimator)
        cam = camera.to_abi()
        out = np.zeros((height, width, 3), dtype=np.float32)
        st = _abi.RenderStats()
        self.ctx.check(self.ctx.lib.ab_render(self.handle, ctypes.byref(cam), ctypes.byref(p), out.ctypes.data,
                                              ctypes.byref(st)))
        if stats:
            return out, {"camera_rays": st.camera_rays, "extension_rays": st.extension_rays,
                         "shadow_rays": st.shadow_rays, "paths": st.paths, "gpu_ms": st.gpu_ms}
        return out


    def render_sums(self, camera: CameraSettings, width, height, samples_per_axis, max_depth=64, roulette_start=3,
                    seed=0, estimator=_abi.ESTIMATOR_PATH, into=None):
        """Adds the render's per-pixel sums to `into` (uint64 [height, width, 4],
        the wire format; zeros when None) and returns it (ab_render_sums: the
        reference renderImage's accumulating form)."""
        from .workloads import render_params
        p = render_params(width, height, samples_per_axis, max_depth, roulette_start, seed, estimator)
        cam = camera.to_abi()
        if into is None:
            into = np.zeros((height, width, 4), np.uint64)
        if into.dtype != np.uint64 or not into.flags.c_contiguous or into.shape != (height, width, 4):
            raise ValueError("render_sums needs a C-contiguous uint64 [height, width, 4] array")
        self.ctx.check(self.ctx.lib.ab_render_sums(self.handle, ctypes.byref(cam), ctypes.byref(p), into.ctypes.data,
                                                   None))
        return into


def hits_t(hits: np.ndarray) -> np.ndarray:
    """The t column of a hits array as float32."""
    return hits[:, 0].copy().view(np.float32)


class Server:
    """The serving loop of a context (ab_server_start): submit jobs, poll
    progressive results, wait, cancel. Jobs are (GpuScene, CameraSettings,
    width, height, samples_per_axis, max_depth, roulette_start, seed)."""

    QUEUED, RUNNING, DONE, CANCELLED, FAILED = 0, 1, 2, 3, 4

    def __init__(self, ctx: Context, hbm_budget_bytes: int, policy=None):
        self.ctx = ctx
        ctx.check(ctx.lib.ab_server_start(ctx.handle, hbm_budget_bytes))
        self.shapes = {}
        if policy is not None:
            self.set_policy(policy)

    def set_policy(self, policy):
        """_abi.SCHED_FIFO, SCHED_SHARE or SCHED_SMALL_FIRST (the default)."""
        self.ctx.check(self.ctx.lib.ab_server_set_policy(self.ctx.handle, policy))

    def stop(self):
        if self.ctx is not None:
            self.ctx.check(self.ctx.lib.ab_server_stop(self.ctx.handle))
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    @staticmethod
    def job_desc(scene: GpuScene, camera: CameraSettings, width, height, samples_per_axis, max_depth=64,
                 roulette_start=3, seed=0):
        from .workloads import render_params
        d = _abi.JobDesc()
        d.scene = scene.handle
        d.camera = camera.to_abi()
        d.params = render_params(width, height, samples_per_axis, max_depth, roulette_start, seed)
        return d

    def submit(self, scene, camera, width, height, samples_per_axis, max_depth=64, roulette_start=3, seed=0,
               keep_sums=False):
        """keep_sums: the finished job still answers poll_sums (AB_JOB_KEEP_SUMS)."""
        d = self.job_desc(scene, camera, width, height, samples_per_axis, max_depth, roulette_start, seed)
        d.flags = _abi.JOB_KEEP_SUMS if keep_sums else 0
        jid = ctypes.c_uint64()
        self.ctx.check(self.ctx.lib.ab_job_submit(self.ctx.handle, ctypes.byref(d), ctypes.byref(jid)))
        self.shapes[jid.value] = (height, width)
        return jid.value

    @staticmethod
    def footprint(scene, camera, width, height, samples_per_axis, max_depth=64, roulette_start=3, seed=0):
        """HBM bytes the job holds while admitted (what the budget is charged)."""
        d = Server.job_desc(scene, camera, width, height, samples_per_axis, max_depth, roulette_start, seed)
        b = ctypes.c_uint64()
        check(lib().ab_job_footprint(ctypes.byref(d), ctypes.byref(b)))
        return b.value

    def _status(self, st):
        return {"state": st.state, "samples_done": st.samples_done, "samples_total": st.samples_total,
                "submit_ms": st.submit_ms, "start_ms": st.start_ms, "finish_ms": st.finish_ms}

    def poll_sums(self, jid):
        """(status code, status dict, sums uint64 [h, w, 4]): the wire format of
        ab_job_poll_sums (red, green, blue fixed-point sums and the sample count)."""
        h, w = self.shapes[jid]
        out = np.zeros((h, w, 4), np.uint64)
        st = _abi.JobStatus()
        code = self.ctx.lib.ab_job_poll_sums(self.ctx.handle, jid, ctypes.byref(st), out.ctypes.data)
        if code not in (_abi.AB_OK, _abi.AB_PENDING, _abi.AB_CANCELLED):
            self.ctx.check(code)
        return code, self._status(st), out

    def poll(self, jid, image=True):
        """(status code, status dict, image or None)."""
        if jid not in self.shapes:
            image = False
        h, w = self.shapes.get(jid, (0, 0))
        out = np.zeros((h, w, 3), np.float32) if image else None
        st = _abi.JobStatus()
        code = self.ctx.lib.ab_job_poll(self.ctx.handle, jid, ctypes.byref(st),
                                        out.ctypes.data if image else None)
        if code not in (_abi.AB_OK, _abi.AB_PENDING, _abi.AB_CANCELLED):
            self.ctx.check(code)
        return code, self._status(st), out

    def wait(self, jid):
        h, w = self.shapes[jid]
        out = np.zeros((h, w, 3), np.float32)
        st = _abi.JobStatus()
        code = self.ctx.lib.ab_job_wait(self.ctx.handle, jid, ctypes.byref(st), out.ctypes.data)
        if code not in (_abi.AB_OK, _abi.AB_CANCELLED):
            self.ctx.check(code)
        return code, self._status(st), out

    def cancel(self, jid):
        self.ctx.check(self.ctx.lib.ab_job_cancel(self.ctx.handle, jid))

    def release(self, jid):
        self.ctx.check(self.ctx.lib.ab_job_release(self.ctx.handle, jid))
        self.shapes.pop(jid, None)

    def stats(self):
        s = _abi.ServerStats()
        self.ctx.check(self.ctx.lib.ab_server_stats_get(self.ctx.handle, ctypes.byref(s)))
        return {"jobs_done": s.jobs_done, "jobs_rejected": s.jobs_rejected, "rays": s.rays, "busy_ms": s.busy_ms,
                "p50_latency_ms": s.p50_latency_ms, "p99_latency_ms": s.p99_latency_ms,
                "hbm_charged": s.hbm_charged}


def sums_resolve(sums: np.ndarray) -> np.ndarray:
    """Per-pixel means of wire-format sums [..., 4] (ab_sums_resolve), float32 [..., 3]."""
    s = np.ascontiguousarray(sums, dtype=np.uint64)
    npx = s.size // 4
    out = np.zeros(s.shape[:-1] + (3,), np.float32)
    check(lib().ab_sums_resolve(s.ctypes.data, npx, out.ctypes.data))
    return out


def sums_merge(dst: np.ndarray, src: np.ndarray) -> np.ndarray:
    """dst += src over wire-format sums (ab_sums_merge), in place; returns dst."""
    if dst.dtype != np.uint64 or not dst.flags.c_contiguous or dst.shape != src.shape:
        raise ValueError("sums_merge needs two C-contiguous uint64 arrays of one shape")
    s = np.ascontiguousarray(src, dtype=np.uint64)
    check(lib().ab_sums_merge(dst.ctypes.data, s.ctypes.data, dst.size // 4))
    return dst


def render_batch(ctx: Context, jobs):
    """jobs: list of (GpuScene, CameraSettings, width, height, spa, max_depth, rr, seed).
    Returns (images, stats dict) from one ab_render_batch call."""
    descs = (_abi.JobDesc * len(jobs))()
    outs = []
    for i, j in enumerate(jobs):
        descs[i] = Server.job_desc(*j)
        outs.append(np.zeros((j[3], j[2], 3), np.float32))
    ptrs = (ctypes.c_void_p * len(jobs))(*[o.ctypes.data for o in outs])
    st = _abi.RenderStats()
    ctx.check(ctx.lib.ab_render_batch(ctx.handle, descs, len(jobs), ptrs, ctypes.byref(st)))
    return outs, {"camera_rays": st.camera_rays, "extension_rays": st.extension_rays, "shadow_rays": st.shadow_rays,
                  "paths": st.paths, "gpu_ms": st.gpu_ms}
