"""Pins the oracle: the reference's own golden vectors and known-answer tests,
closed forms, and brute force. Each test names the reference test it follows.
A test that plants a mistake plants it in a copy of the oracle's formula here,
on the CPU, to show the check would catch it."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from oracle import dvec

L = oracle.lib()


# ------------------------------------------------------------------ sampling

def pcg(seed, stream, n, advance=0):
    out = (ctypes.c_uint32 * n)()
    L.or_pcg32(seed, stream, advance, n, out)
    return list(out)


def test_pcg32_matches_reference_demo_output():
    # tests/sampling/pcg32_test.cpp:12-23 (pcg32-demo.c, seed 42, stream 54)
    assert pcg(42, 54, 6) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_advance_equals_stepping():
    seq = pcg(7, 3, 3000)
    for k in (0, 1, 2, 17, 1024, 2999):
        assert pcg(7, 3, 1, advance=k)[0] == seq[k]


def test_pcg32_chi_square():
    # tests/sampling/pcg32_test.cpp:64-83: 16 buckets, 10 000 draws.
    draws = np.array(pcg(1234, 5, 16000), dtype=np.uint64)
    counts = np.bincount((draws >> np.uint64(28)).astype(np.int64), minlength=16)
    expected = len(draws) / 16
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 37.7  # 99.9th percentile of chi-square with 15 dof


def test_mixbits_is_splitmix64():
    # Vigna's splitmix64.c with state 0: the first output is the finaliser of
    # the golden-ratio increment.
    assert L.or_mixbits(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    assert L.or_mixbits(0) == 0


def test_vectorised_pcg_workload_generator_matches_oracle():
    from albedo_amd.workloads import pcg32_stream
    got = pcg32_stream(0x5EED, 1, 10000)
    ref = np.array(pcg(0x5EED, 1, 10000), dtype=np.uint32)
    assert np.array_equal(got, ref)


# ---------------------------------------------------------------- textures

@pytest.mark.parametrize("p,expected", [((3.14, 42, 7), 0.13691995878400012),
                                        ((-0.3, -5.7, 2.25), 0.40750939192218755),
                                        ((-300.6, 17.2, -0.45), -0.11245507194489657)])
def test_perlin_matches_improved_noise(p, expected):
    # tests/texture/perlin_noise_test.cpp:79-87
    assert abs(L.or_perlin(*p) - expected) < 1e-12


def test_perlin_zero_on_lattice_and_periodic():
    # perlin_noise_test.cpp: zero at integer points, period 256.
    for p in [(0, 0, 0), (3, -7, 12), (-255, 1, 2)]:
        assert L.or_perlin(*p) == 0.0
    a = L.or_perlin(0.42, 1.37, -2.71)
    b = L.or_perlin(0.42 + 256, 1.37 - 512, -2.71 + 256)
    assert abs(a - b) < 1e-9


def test_turbulence_is_seven_octave_sum():
    p = (0.3, 1.7, -2.2)
    s, w, q = 0.0, 1.0, p
    for _ in range(7):
        s += w * abs(L.or_perlin(*q))
        w *= 0.5
        q = tuple(2 * c for c in q)
    assert abs(L.or_turbulence(*p) - s) < 1e-15
    # marble: (1 + sin(scale x + veins turbulence)) / 2 (perlin_pattern.cpp:11-23)
    scale, veins = 2.0, 5.0
    sp = tuple(scale * c for c in p)
    expect = (1 + math.sin(sp[0] + veins * L.or_turbulence(*sp))) / 2
    assert abs(L.or_pattern(*p, scale, veins) - expect) < 1e-15
    assert L.or_pattern(*p, scale, 0.0) == min(L.or_turbulence(*sp), 1.0)


# ------------------------------------------------------------------ geometry

def test_sphere_near_root_matches_60_digit_value():
    # tests/geometry/sphere_test.cpp:225-240
    t = ctypes.c_double()
    ok = L.or_sphere_t(dvec(0, 0, 0), 1.0, dvec(0.2957763671875, 0.15106201171875, 0.9432373046875),
                       dvec(-0.3411177909433509, -0.17474016395585198, -0.9432373046875), 0.0, math.inf,
                       ctypes.byref(t))
    assert ok
    exact = 1.8315314704767794e-09
    assert abs(t.value - exact) <= exact * 1e-12


def test_sphere_textbook_root_would_fail_that_check():
    # The planted mistake: (-h - sqrt(disc)) / a cancels and misses the bound.
    o = np.array([0.2957763671875, 0.15106201171875, 0.9432373046875])
    d = np.array([-0.3411177909433509, -0.17474016395585198, -0.9432373046875])
    a, h, c = d @ d, d @ o, o @ o - 1
    t_bad = (-h - math.sqrt(h * h - a * c)) / a
    assert abs(t_bad - 1.8315314704767794e-09) > 1.8315314704767794e-09 * 1e-12


def sphere_t(c, r, o, d, tmin=0.0, tmax=math.inf):
    t = ctypes.c_double()
    ok = L.or_sphere_t(dvec(*c), r, dvec(*o), dvec(*d), tmin, tmax, ctypes.byref(t))
    return t.value if ok else None


def test_sphere_hits_near_then_far_and_respects_range():
    assert abs(sphere_t((0, 0, -5), 1, (0, 0, 0), (0, 0, -1)) - 4) < 1e-12
    assert abs(sphere_t((0, 0, -5), 1, (0, 0, 0), (0, 0, -1), 4.5) - 6) < 1e-12
    assert sphere_t((0, 0, -5), 1, (0, 0, 0), (0, 0, -1), 0, 4.0) is None  # surrounds() is strict
    assert abs(sphere_t((0, 0, 0), 1, (0, 0, 0), (0, 2, 0)) - 0.5) < 1e-12  # unnormalised direction
    assert sphere_t((0, 0, -5), 1, (0, 0, 0), (0, 1, 0)) is None


def tri(v, o, d, fp32, tmin=0.0, tmax=math.inf):
    if fp32:
        vv = (ctypes.c_float * 9)(*v)
        out = (ctypes.c_float * 3)()
        ok = L.or_triangle_f(vv, (ctypes.c_float * 3)(*o), (ctypes.c_float * 3)(*d), tmin, tmax, out)
    else:
        vv = (ctypes.c_double * 9)(*v)
        out = (ctypes.c_double * 3)()
        ok = L.or_triangle_d(vv, dvec(*o), dvec(*d), tmin, tmax, out)
    return tuple(out) if ok else None


@pytest.mark.parametrize("fp32", [False, True])
def test_triangle_every_dominant_axis(fp32):
    # tests/geometry/triangle_test.cpp:268-289
    v = (1, 0, 0, 0, 1, 0, 0, 0, 1)
    target = np.array([1 / 3, 1 / 3, 1 / 3])
    for off in [(3, .2, .1), (-3, .2, .1), (.1, 3, .2), (.1, -3, .2), (.2, .1, 3), (.2, .1, -3)]:
        off = np.array(off)
        h = tri(v, target + off, -off, fp32)
        assert h is not None and abs(h[0] - 1) < (1e-6 if fp32 else 1e-12)


@pytest.mark.parametrize("fp32", [False, True])
def test_triangle_shared_edge_leaves_no_gap(fp32):
    # tests/geometry/triangle_test.cpp:291-322
    a, b, c, d = (np.array(p) for p in [(-1.3, -.9, -4.7), (1.1, -1.2, -5.9), (1.4, 1.0, -5.3), (-.8, 1.3, -4.1)])
    if fp32:
        a, b, c, d = (p.astype(np.float32).astype(np.float64) for p in (a, b, c, d))
    first = tuple(np.concatenate([a, b, c]))
    second = tuple(np.concatenate([a, c, d]))
    misses = 0
    for origin in [(.123, -.271, .317), (.217, .131, -12.71), (21.3, .311, -5.07), (.173, -19.7, -4.93)]:
        o = np.array(origin)
        for i in range(1000):
            s = (i + 0.5) / 1000
            e = a + (c - a) * s
            dd = e - o
            if tri(first, o, dd, fp32) is None and tri(second, o, dd, fp32) is None:
                misses += 1
    assert misses == 0


@pytest.mark.parametrize("fp32", [False, True])
def test_ray_exactly_on_shared_edge_hits_both(fp32):
    # tests/geometry/triangle_test.cpp:324-335 (edge function exactly zero; in
    # FP32 this is where the FP64 recomputation runs)
    above = (-1, 0, -5, 1, 0, -5, 0, 1, -5)
    below = (1, 0, -5, -1, 0, -5, 0, -1, -5)
    o, d = (0.25, 0, 0), (0, 0, -1)
    assert tri(above, o, d, fp32) is not None
    assert tri(below, o, d, fp32) is not None


def test_fp32_triangle_agrees_with_fp64_on_random_hits():
    rng = np.random.default_rng(3)
    agree = 0
    for _ in range(2000):
        v = rng.uniform(-1, 1, 9).astype(np.float32)
        o = (rng.uniform(-3, 3, 3)).astype(np.float32)
        tgt = (v[0:3] + v[3:6] + v[6:9]) / 3 + rng.uniform(-.3, .3, 3).astype(np.float32)
        d = (tgt - o).astype(np.float32)
        h32 = tri(tuple(float(x) for x in v), o, d, True)
        h64 = tri(tuple(float(x) for x in v), o.astype(np.float64), d.astype(np.float64), False)
        if h32 is None or h64 is None:
            continue
        agree += 1
        assert abs(h32[0] - h64[0]) <= 64 * np.spacing(np.float32(h64[0])) + 1e-6 * abs(h64[0])
    assert agree > 500


# ------------------------------------------------------------------ fresnel

def test_conductor_fresnel_at_normal_incidence_is_closed_form():
    # tests/scattering/fresnel_test.cpp:79-90
    eta, k = (0.2, 0.9, 1.4), (3.0, 2.4, 1.6)
    out = (ctypes.c_double * 3)()
    L.or_fresnel_conductor(1.0, dvec(*eta), dvec(*k), out)
    for i in range(3):
        n, kk = eta[i], k[i]
        assert abs(out[i] - ((n - 1) ** 2 + kk ** 2) / ((n + 1) ** 2 + kk ** 2)) < 1e-12


def test_conductor_without_absorption_is_dielectric():
    # tests/scattering/fresnel_test.cpp:92-100
    out = (ctypes.c_double * 3)()
    for c in (1.0, 0.8, 0.4, 0.1):
        L.or_fresnel_conductor(c, dvec(1.5, 1.5, 1.5), dvec(0, 0, 0), out)
        assert abs(out[0] - L.or_fresnel_dielectric(c, 1.5)) < 1e-12


def test_dielectric_fresnel_closed_forms():
    assert abs(L.or_fresnel_dielectric(1.0, 1.5) - 0.04) < 1e-12
    assert L.or_fresnel_dielectric(math.cos(0.8), 1 / 1.5) == 1.0  # total internal reflection
    assert abs(L.or_fresnel_dielectric(-0.6, 1.5) - L.or_fresnel_dielectric(0.6, 1 / 1.5)) < 1e-12
    assert L.or_fresnel_dielectric(0.3, 1.0) == 0.0
    # Against Schlick-free textbook s/p form at 45 degrees, eta 1.5.
    ci = math.cos(math.pi / 4)
    st = math.sin(math.pi / 4) / 1.5
    ct = math.sqrt(1 - st * st)
    rs = ((ci - 1.5 * ct) / (ci + 1.5 * ct)) ** 2
    rp = ((1.5 * ci - ct) / (1.5 * ci + ct)) ** 2
    assert abs(L.or_fresnel_dielectric(ci, 1.5) - (rs + rp) / 2) < 1e-12


# ---------------------------------------------------------------------- GGX

def test_ggx_density_and_masking_closed_forms():
    # tests/scattering/ggx_distribution_test.cpp:54-59, 85-91
    assert abs(L.or_ggx_density(0.3, dvec(0, 0, 1)) - 1 / (math.pi * 0.09)) < 1e-9
    th, ph = math.pi / 3, 0.4
    w = dvec(math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
    assert abs(L.or_ggx_masking(1.0, w) - 2 / 3) < 1e-12
    assert L.or_ggx_density(0.3, dvec(0, 0.6, -0.8)) == 0.0


def test_ggx_projected_density_integrates_to_one():
    # ggx_distribution_test.cpp:62-70 (midpoint rule on the hemisphere)
    for alpha in (0.2, 0.5, 1.0):
        n_t, n_p = 800, 64
        s = 0.0
        for i in range(n_t):
            th = (i + 0.5) * (math.pi / 2) / n_t
            m = dvec(math.sin(th), 0, math.cos(th))
            s += L.or_ggx_density(alpha, m) * math.cos(th) * math.sin(th)
        s *= (math.pi / 2) / n_t * 2 * math.pi
        assert abs(s - 1) < 2e-3


# -------------------------------------------------------------------- lobes

def lobe_sample(kind, p, wo, uc, ux, uy):
    wi, wt, pdf, spec = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), ctypes.c_double(), ctypes.c_int()
    ok = L.or_lobe_sample(kind, dvec(*p), dvec(*wo), uc, ux, uy, wi, wt, ctypes.byref(pdf), ctypes.byref(spec))
    return (np.array(wi[:]), np.array(wt[:]), pdf.value, bool(spec.value)) if ok else None


def lobe_eval(kind, p, wo, wi):
    f, pdf = (ctypes.c_double * 3)(), ctypes.c_double()
    L.or_lobe_eval(kind, dvec(*p), dvec(*wo), dvec(*wi), f, ctypes.byref(pdf))
    return np.array(f[:]), pdf.value


LOBES = [(0, (0.5, 0.6, 0.7)),
         (1, (0.156, 0.424, 1.383, 3.602, 2.472, 1.915, 0.0625)),
         (2, (1.5, 0.04)),
         (2, (1 / 1.5, 0.09))]


@pytest.mark.parametrize("kind,p", LOBES)
def test_lobe_sample_weight_matches_eval_over_pdf(kind, p):
    # tests/scattering/*_lobe_test.cpp: samples agree with evaluate and pdf.
    rng = np.random.default_rng(kind)
    wo = np.array([0.3, -0.2, 0.0])
    wo[2] = math.sqrt(1 - wo @ wo)
    checked = 0
    for _ in range(400):
        s = lobe_sample(kind, p, wo, *rng.uniform(0, 1, 3))
        if s is None or s[3]:
            continue
        wi, wt, pdf, _ = s
        f, pdf2 = lobe_eval(kind, p, wo, wi)
        assert abs(pdf - pdf2) <= 1e-9 * max(1, pdf)
        np.testing.assert_allclose(wt, f * abs(wi[2]) / pdf, rtol=1e-9, atol=1e-12)
        checked += 1
    assert checked > 100


ROUGH_LOBES = [(0, (0.5, 0.6, 0.7)),
               (1, (0.156, 0.424, 1.383, 3.602, 2.472, 1.915, 0.5)),
               (2, (1.5, 0.4)),
               (2, (1 / 1.5, 0.3))]


@pytest.mark.parametrize("kind,p", ROUGH_LOBES)
def test_lobe_pdf_integrates_to_at_most_one(kind, p):
    # Uniform-sphere Monte Carlo of the pdf; rough lobes keep its variance low.
    wo = np.array([0.3, -0.2, 0.0])
    wo[2] = math.sqrt(1 - wo @ wo)
    rng = np.random.default_rng(11)
    n = 200000
    u = rng.uniform(0, 1, (n, 2))
    z = 1 - 2 * u[:, 0]
    r = np.sqrt(np.maximum(0, 1 - z * z))
    ph = 2 * math.pi * u[:, 1]
    dirs = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    total = 0.0
    for d in dirs[:100000]:
        total += lobe_eval(kind, p, wo, d)[1]
    est = total / 100000 * 4 * math.pi
    assert est < 1.05
    assert est > (0.9 if kind == 0 else 0.5)


def test_lambert_eval_is_albedo_over_pi():
    f, pdf = lobe_eval(0, (0.5, 0.6, 0.7), (0, 0, 1), (0.6, 0, 0.8))
    np.testing.assert_allclose(f, np.array([0.5, 0.6, 0.7]) / math.pi, rtol=1e-15)
    assert abs(pdf - 0.8 / math.pi) < 1e-15


def test_mirror_and_smooth_glass_are_specular():
    wo = (0.6, 0.0, 0.8)
    s = lobe_sample(1, (0.155, 0.117, 0.138, 4.83, 3.12, 2.15, 0.0), wo, 0.5, 0.5, 0.5)
    assert s[3] and np.allclose(s[0], (-0.6, 0.0, 0.8))
    refl = lobe_sample(2, (1.5, 0.0), wo, 0.0, 0.5, 0.5)
    refr = lobe_sample(2, (1.5, 0.0), wo, 0.999, 0.5, 0.5)
    assert refl[3] and refr[3]
    assert refl[0][2] > 0 and refr[0][2] < 0
    assert abs(refr[1][0] - 1 / 1.5 ** 2) < 1e-15
    # Snell: sin_t = sin_i / eta
    assert abs(math.hypot(refr[0][0], refr[0][1]) - 0.6 / 1.5) < 1e-12


# ------------------------------------------------------ estimators (closed forms)

def _one_sphere_scene(sky=None, albedo=0.6):
    from albedo_amd.scene import ConstantTexture, DiffuseMaterial, Scene, SkyLight, Sphere
    s = Scene()
    s.add(Sphere((0.0, 0.0, 0.0), 1.0), DiffuseMaterial(ConstantTexture((albedo, albedo, albedo))))
    if sky is not None:
        s.set_sky(SkyLight.uniform(sky))
    return oracle.OracleScene(s.to_desc())


def _narrow_camera(look_at):
    from albedo_amd import _abi
    c = _abi.Camera()
    c.look_from[:] = (0.0, 0.0, 5.0)
    c.look_at[:] = look_at
    c.up[:] = (0.0, 1.0, 0.0)
    c.vfov_degrees = 0.001  # a pixel spans 4e-5 at the sphere
    c.aspect = 1.0
    return c


def test_normals_view_closed_forms():
    # Reference normals_estimator.cpp:15-40: a hit shows 0.5 (n + 1); a miss the
    # white-to-blue gradient on the direction's height.
    from albedo_amd import _abi
    from albedo_amd import workloads as W
    o = _one_sphere_scene()
    p = W.render_params(1, 1, 2, 8, 3, 0, _abi.ESTIMATOR_NORMALS)
    img, counts = o.render(_narrow_camera((0.0, 0.0, 0.0)), p, sample_mode=0)
    assert np.allclose(img[0, 0], [0.5, 0.5, 1.0], atol=1e-4)
    assert counts[1] == 0 and counts[2] == 0
    img, _ = o.render(_narrow_camera((0.0, 0.0, 10.0)), p, sample_mode=0)  # looking away: a miss along +z
    assert np.allclose(img[0, 0], [0.75, 0.85, 1.0], atol=1e-4)


def test_scatter_only_convex_object_under_uniform_sky_is_exact():
    # Reference scatter_only_estimator.cpp:10-48: a ray that hits a lone convex
    # diffuse object scatters once (weight = albedo exactly, lambert_lobe.cpp) and
    # then escapes, so every sample is albedo x sky radiance.
    from albedo_amd import _abi
    from albedo_amd import workloads as W
    o = _one_sphere_scene(sky=(0.5, 0.5, 0.5), albedo=0.6)
    p = W.render_params(1, 1, 8, 32, 3, 4, _abi.ESTIMATOR_SCATTER_ONLY)
    img, counts = o.render(_narrow_camera((0.0, 0.0, 0.0)), p, sample_mode=0)
    exact = float(np.float32(0.6)) * float(np.float32(0.5))  # the scene stores colours in FP32
    assert np.allclose(img[0, 0], exact, rtol=1e-12)
    assert counts[0] == 64 and counts[1] == 64 and counts[2] == 0
    # A planted mistake: dropping the sky term at the escape would give black.
    p.max_depth = 0  # scatter limit 0: the path ends at its first hit, gathering nothing
    img, _ = o.render(_narrow_camera((0.0, 0.0, 0.0)), p, sample_mode=0)
    assert np.allclose(img[0, 0], 0.0)
