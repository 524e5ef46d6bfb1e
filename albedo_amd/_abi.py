"""ctypes view of include/albedo_gpu.h and the loader of libalbedo_gpu.so.

The structures here mirror the C header field for field; tests check their sizes
against the compiled library's expectations. Loading fails loudly when the
library has not been built: there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libalbedo_gpu.so")

c_float3 = ctypes.c_float * 3
c_double3 = ctypes.c_double * 3

AB_OK = 0
AB_INVALID_ARGUMENT = 1
AB_OUT_OF_MEMORY = 2
AB_DEVICE_ERROR = 3
AB_REJECTED_BUDGET = 4
AB_NOT_FOUND = 5
AB_CANCELLED = 6
AB_PENDING = 7
AB_PARSE_ERROR = 8
STATUS_NAMES = {0: "AB_OK", 1: "AB_INVALID_ARGUMENT", 2: "AB_OUT_OF_MEMORY", 3: "AB_DEVICE_ERROR",
                4: "AB_REJECTED_BUDGET", 5: "AB_NOT_FOUND", 6: "AB_CANCELLED", 7: "AB_PENDING",
                8: "AB_PARSE_ERROR"}

AB_TUNE_PLAIN, AB_TUNE_BATCH = 0, 1
ESTIMATOR_PATH, ESTIMATOR_SCATTER_ONLY, ESTIMATOR_NORMALS = 0, 1, 2
SCHED_FIFO, SCHED_SHARE, SCHED_SMALL_FIRST = 0, 1, 2
JOB_KEEP_SUMS = 1
BUILD_DEVICE_BLAS = 1
BUILD_WIDE_BVH = 2
BUILD_ONE_PASS = 4
SUMS_FRACTION_BITS = 28

AB_MISS = 0xFFFFFFFF
AB_NO_PRIM = 0xFFFFFFFF

TEX_CONSTANT, TEX_CHECKER, TEX_PERLIN = 0, 1, 2
MAT_DIFFUSE, MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_MIX = 0, 1, 2, 3
OBJ_SPHERE, OBJ_INSTANCE = 0, 1
LIGHT_POINT, LIGHT_AREA, LIGHT_SKY = 0, 1, 2


class Ray(ctypes.Structure):
    _fields_ = [("org", c_float3), ("tmin", ctypes.c_float), ("dir", c_float3), ("tmax", ctypes.c_float)]


class Hit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("object", ctypes.c_uint32), ("prim", ctypes.c_uint32),
                ("bary", ctypes.c_uint32)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("positions", ctypes.POINTER(ctypes.c_float)), ("num_vertices", ctypes.c_uint32),
                ("indices", ctypes.POINTER(ctypes.c_uint32)), ("num_faces", ctypes.c_uint32),
                ("uv", ctypes.POINTER(ctypes.c_float)), ("smooth", ctypes.c_int32)]


class TextureDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("a", c_float3), ("b", c_float3), ("frequency", ctypes.c_float),
                ("veins", ctypes.c_float)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("albedo", TextureDesc), ("eta", c_float3), ("k", c_float3),
                ("roughness", ctypes.c_float), ("ior", ctypes.c_float), ("first", ctypes.c_uint32),
                ("second", ctypes.c_uint32), ("scale", ctypes.c_float), ("veins", ctypes.c_float),
                ("threshold", ctypes.c_float)]


class ObjectDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("material", ctypes.c_uint32), ("light", ctypes.c_int32),
                ("mesh", ctypes.c_uint32), ("center", c_double3), ("radius", ctypes.c_double),
                ("to_world", ctypes.c_double * 12), ("to_object", ctypes.c_double * 12)]


class LightDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("object", ctypes.c_uint32), ("position", c_double3),
                ("intensity", c_float3), ("radiance", c_float3), ("nadir", c_float3), ("zenith", c_float3)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("meshes", ctypes.POINTER(MeshDesc)), ("num_meshes", ctypes.c_uint32),
                ("objects", ctypes.POINTER(ObjectDesc)), ("num_objects", ctypes.c_uint32),
                ("materials", ctypes.POINTER(MaterialDesc)), ("num_materials", ctypes.c_uint32),
                ("lights", ctypes.POINTER(LightDesc)), ("num_lights", ctypes.c_uint32),
                ("build_flags", ctypes.c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [("look_from", c_double3), ("look_at", c_double3), ("up", c_double3),
                ("vfov_degrees", ctypes.c_double), ("aspect", ctypes.c_double)]


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("samples_per_axis", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("roulette_start", ctypes.c_uint32), ("estimator", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class RenderStats(ctypes.Structure):
    _fields_ = [("camera_rays", ctypes.c_uint64), ("extension_rays", ctypes.c_uint64),
                ("shadow_rays", ctypes.c_uint64), ("paths", ctypes.c_uint64), ("gpu_ms", ctypes.c_double)]


class LaunchConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("paths_in_flight", ctypes.c_uint32),
                ("grid_blocks", ctypes.c_uint32), ("use_graph", ctypes.c_uint32)]


class JobDesc(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("camera", Camera), ("params", RenderParams), ("flags", ctypes.c_uint32)]


class JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("samples_done", ctypes.c_uint64), ("samples_total", ctypes.c_uint64),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("finish_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("jobs_done", ctypes.c_uint64), ("jobs_rejected", ctypes.c_uint64), ("rays", ctypes.c_uint64),
                ("busy_ms", ctypes.c_double), ("p50_latency_ms", ctypes.c_double),
                ("p99_latency_ms", ctypes.c_double), ("hbm_charged", ctypes.c_uint64)]


# Every entry point of include/albedo_gpu.h: name -> (restype, argtypes).
_P = ctypes.c_void_p
_S = ctypes.c_int
PROTOTYPES = {
    "ab_abi_version": (ctypes.c_int, []),
    "ab_context_create": (_S, [ctypes.c_int, ctypes.POINTER(_P)]),
    "ab_context_destroy": (None, [_P]),
    "ab_last_error": (ctypes.c_char_p, [_P]),
    "ab_scene_create": (_S, [_P, ctypes.POINTER(SceneDesc), ctypes.POINTER(_P)]),
    "ab_scene_release": (None, [_P]),
    "ab_scene_hash": (ctypes.c_uint64, [_P]),
    "ab_scene_device_bytes": (ctypes.c_uint64, [_P]),
    "ab_scene_info": (_S, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64),
                           ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "ab_trace_closest": (_S, [_P, _P, _P, ctypes.c_uint64, _P]),
    "ab_trace_any": (_S, [_P, _P, _P, ctypes.c_uint64, _P]),
    "ab_trace_closest_host": (_S, [_P, _P, _P, ctypes.c_uint64]),
    "ab_trace_any_host": (_S, [_P, _P, _P, ctypes.c_uint64]),
    "ab_render": (_S, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P, ctypes.POINTER(RenderStats)]),
    "ab_render_sums": (_S, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P,
                            ctypes.POINTER(RenderStats)]),
    "ab_render_static": (_S, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P, _P]),
    "ab_set_launch_config": (_S, [_P, ctypes.POINTER(LaunchConfig)]),
    "ab_get_launch_config": (_S, [_P, ctypes.POINTER(LaunchConfig)]),
    "ab_render_batch": (_S, [_P, ctypes.POINTER(JobDesc), ctypes.c_uint32, ctypes.POINTER(_P),
                             ctypes.POINTER(RenderStats)]),
    "ab_autotune": (_S, [_P, ctypes.POINTER(JobDesc), ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(LaunchConfig),
                         ctypes.POINTER(ctypes.c_double)]),
    "ab_server_start": (_S, [_P, ctypes.c_uint64]),
    "ab_server_stop": (_S, [_P]),
    "ab_server_set_policy": (_S, [_P, ctypes.c_uint32]),
    "ab_job_submit": (_S, [_P, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_uint64)]),
    "ab_job_footprint": (_S, [ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_uint64)]),
    "ab_job_poll": (_S, [_P, ctypes.c_uint64, ctypes.POINTER(JobStatus), _P]),
    "ab_job_wait": (_S, [_P, ctypes.c_uint64, ctypes.POINTER(JobStatus), _P]),
    "ab_job_poll_sums": (_S, [_P, ctypes.c_uint64, ctypes.POINTER(JobStatus), _P]),
    "ab_sums_merge": (_S, [_P, _P, ctypes.c_uint64]),
    "ab_sums_resolve": (_S, [_P, ctypes.c_uint64, _P]),
    "ab_job_cancel": (_S, [_P, ctypes.c_uint64]),
    "ab_job_release": (_S, [_P, ctypes.c_uint64]),
    "ab_server_stats_get": (_S, [_P, ctypes.POINTER(ServerStats)]),
}

_lib = None


def lib():
    """The loaded library with prototypes set. Raises if it is not built."""
    global _lib
    if _lib is None:
        # PyTorch ships its own HIP runtime under the same soname. Whichever is
        # loaded first serves the whole process, and torch only initialises on its
        # own copy, so load torch (when installed) before the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libalbedo_gpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class AlbedoError(RuntimeError):
    """A non-OK status from the library, with its message."""

    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


def check(status, ctx=None):
    if status != AB_OK:
        msg = lib().ab_last_error(ctx)
        raise AlbedoError(status, msg.decode() if msg else "")
