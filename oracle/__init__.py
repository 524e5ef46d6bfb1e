"""The oracle: a CPU restatement of the reference renderer (anazli/albedo), used
only as the checker by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg. The product library never imports, links or calls it.

liboracle.so (built from oracle_api.cpp by oracle/Makefile) exposes:
  trace_closest / trace_any   FP64 restatement (mode 0) or FP32 mirror of the
                              GPU's operation order (mode 1); BVH or brute force
  render                      FP64 path tracer, reference draw order (sample_mode
                              0) or the GPU's per-sample streams (sample_mode 1)
  unit functions              PCG32, SplitMix, Perlin, Fresnel, GGX, lobes,
                              sphere and triangle tests

Each function's reference source is cited in oracle_core.hpp, oracle_shade.hpp
and oracle_api.cpp. Parity pins (golden vectors, closed forms, brute force) are
in tests/test_oracle_*.py.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.c_uint64


def build():
    """Compiles liboracle.so with the committed Makefile."""
    r = subprocess.run(["make", "-s", "-C", _HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_scene_create.restype = _P
        L.or_scene_create.argtypes = [_P]
        L.or_scene_destroy.argtypes = [_P]
        L.or_scene_tlas_depth.restype = ctypes.c_uint32
        L.or_scene_tlas_depth.argtypes = [_P]
        L.or_trace_closest.argtypes = [_P, _P, _U64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]
        L.or_trace_any.argtypes = [_P, _P, _U64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        L.or_trace_counts.argtypes = [_P, _P, _U64, ctypes.c_int, _P, _P]
        L.or_render.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, _P, _P]
        L.or_radiance.argtypes = [_P, _DP, _DP, _U64, ctypes.c_int, ctypes.c_int, _U64, _DP]
        L.or_pcg32.argtypes = [_U64, _U64, _U64, ctypes.c_uint32, _P]
        L.or_mixbits.restype = _U64
        L.or_mixbits.argtypes = [_U64]
        for name in ("or_perlin", "or_turbulence"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D, _D, _D]
        L.or_pattern.restype = _D
        L.or_pattern.argtypes = [_D, _D, _D, _D, _D]
        L.or_fresnel_dielectric.restype = _D
        L.or_fresnel_dielectric.argtypes = [_D, _D]
        L.or_fresnel_conductor.argtypes = [_D, _DP, _DP, _DP]
        L.or_sphere_t.restype = ctypes.c_int
        L.or_sphere_t.argtypes = [_DP, _D, _DP, _DP, _D, _D, _DP]
        L.or_triangle_f.restype = ctypes.c_int
        L.or_triangle_f.argtypes = [_P, _P, _P, ctypes.c_float, ctypes.c_float, _P]
        L.or_triangle_d.restype = ctypes.c_int
        L.or_triangle_d.argtypes = [_P, _P, _P, _D, _D, _P]
        L.or_lobe_eval.argtypes = [ctypes.c_int, _DP, _DP, _DP, _DP, _DP]
        L.or_lobe_sample.restype = ctypes.c_int
        L.or_lobe_sample.argtypes = [ctypes.c_int, _DP, _DP, _D, _D, _D, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int)]
        L.or_ggx_density.restype = _D
        L.or_ggx_density.argtypes = [_D, _DP]
        L.or_ggx_masking.restype = _D
        L.or_ggx_masking.argtypes = [_D, _DP]
        L.or_ggx_vn_sample_many.argtypes = [_D, _DP, _U64, _P, _P]
        L.or_ggx_vn_pdf_many.argtypes = [_D, _DP, _U64, _P, _P]
        L.or_lobe_eval_many.argtypes = [ctypes.c_int, _DP, _DP, _U64, _P, _P, _P]
        L.or_lobe_sample_many.restype = _U64
        L.or_lobe_sample_many.argtypes = [ctypes.c_int, _DP, _DP, _U64, _P, _P, _P, _P, _P]
        L.or_radiance_est.argtypes = [_P, _DP, _DP, _U64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _U64, _DP]
        _lib = L
    return _lib


def dvec(*v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


class OracleScene:
    """The oracle's own copy of a scene, built from the same ab_scene_desc."""

    def __init__(self, scene_desc):
        self._desc = scene_desc  # keep the description's buffers alive
        self.h = lib().or_scene_create(ctypes.addressof(scene_desc.desc))
        if not self.h:
            raise RuntimeError("oracle could not build the scene")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    def trace_closest(self, rays, fp32=True, brute=False, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        t = np.zeros(n, np.float64)
        obj = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        b1 = np.zeros(n, np.float64)
        b2 = np.zeros(n, np.float64)
        lib().or_trace_closest(self.h, rays.ctypes.data, n, 1 if fp32 else 0, 1 if brute else 0,
                               threads or os.cpu_count() or 1, t.ctypes.data, obj.ctypes.data, prim.ctypes.data,
                               b1.ctypes.data, b2.ctypes.data)
        return t, obj, prim, b1, b2

    def trace_counts(self, rays, threads=None):
        """(box tests, primitive visits) of the FP64 ordered traversal, summed over rays."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        nodes, prims = ctypes.c_uint64(), ctypes.c_uint64()
        lib().or_trace_counts(self.h, rays.ctypes.data, len(rays), threads or os.cpu_count() or 1,
                              ctypes.byref(nodes), ctypes.byref(prims))
        return nodes.value, prims.value

    def trace_any(self, rays, fp32=True, brute=False, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(len(rays), np.uint8)
        lib().or_trace_any(self.h, rays.ctypes.data, len(rays), 1 if fp32 else 0, 1 if brute else 0,
                           threads or os.cpu_count() or 1, occ.ctypes.data)
        return occ

    def render(self, camera_abi, params_abi, sample_mode=0, threads=None, rows=None):
        """Mean radiance [h, w, 3] (FP64) and ray counts (camera, extension, shadow)."""
        h, w = params_abi.height, params_abi.width
        out = np.zeros((h, w, 3), np.float64)
        counts = np.zeros(3, np.uint64)
        r0, r1 = rows if rows else (0, h)
        lib().or_render(self.h, ctypes.addressof(camera_abi), ctypes.addressof(params_abi), sample_mode,
                        threads or os.cpu_count() or 1, r0, r1, out.ctypes.data, counts.ctypes.data)
        return out, counts

    def radiance(self, origin, direction, seed=17, max_depth=64, roulette_start=3, count=1):
        out = (ctypes.c_double * 3)()
        lib().or_radiance(self.h, dvec(*origin), dvec(*direction), seed, max_depth, roulette_start, count, out)
        return np.array(out[:])

    def radiance_est(self, origin, direction, estimator, seed=17, max_depth=64, roulette_start=3, count=1):
        """Mean of `count` estimates along one ray with the chosen estimator
        (0 path, 1 scatter-only with max_depth as its scatter limit, 2 normals)."""
        out = (ctypes.c_double * 3)()
        lib().or_radiance_est(self.h, dvec(*origin), dvec(*direction), seed, estimator, max_depth, roulette_start,
                              count, out)
        return np.array(out[:])


# ------------------------------------------------------------ array helpers

def _vec3(v):
    return dvec(*[float(x) for x in v])


def ggx_vn_sample(alpha, wo, u):
    """Visible normals sampled from u [n, 2]."""
    u = np.ascontiguousarray(u, np.float64)
    m = np.zeros((len(u), 3))
    lib().or_ggx_vn_sample_many(alpha, _vec3(wo), len(u), u.ctypes.data, m.ctypes.data)
    return m


def ggx_vn_pdf(alpha, wo, m):
    m = np.ascontiguousarray(m, np.float64)
    pdf = np.zeros(len(m))
    lib().or_ggx_vn_pdf_many(alpha, _vec3(wo), len(m), m.ctypes.data, pdf.ctypes.data)
    return pdf


def lobe_eval_many(kind, params, wo, wi):
    """(f [n, 3], pdf [n]) of a lobe (kind 0 Lambert, 1 conductor, 2 dielectric) for wi [n, 3]."""
    wi = np.ascontiguousarray(wi, np.float64)
    f = np.zeros((len(wi), 3))
    pdf = np.zeros(len(wi))
    lib().or_lobe_eval_many(kind, dvec(*params), _vec3(wo), len(wi), wi.ctypes.data, f.ctypes.data, pdf.ctypes.data)
    return f, pdf


def lobe_sample_many(kind, params, wo, u):
    """u [n, 3] = (choice, x, y) -> (wi [n, 3], weight [n, 3], pdf [n], kept [n] bool)."""
    u = np.ascontiguousarray(u, np.float64)
    n = len(u)
    wi, wt, pdf, kept = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.uint8)
    lib().or_lobe_sample_many(kind, dvec(*params), _vec3(wo), n, u.ctypes.data, wi.ctypes.data, wt.ctypes.data,
                              pdf.ctypes.data, kept.ctypes.data)
    return wi, wt, pdf, kept.astype(bool)
