"""Headline benchmark: closest-hit throughput on the trefoil microbenchmark.

Workload: a 512 x 32-segment trefoil tube (32 768 FP32 triangles, binned-SAH
BVH) and two ray sets of 2^24 rays each, coherent (pinhole camera at distance 8)
and incoherent (origins on a radius-6 sphere, targets uniform in the mesh box).
One step traces both sets once with ab_trace_closest, rays and hits resident in
HBM. value = rays / second over the timed steps, in Mrays/s.

roofline: 48 algorithmic bytes per ray (32 B ray read + 16 B hit written) over
the closest-hit kernel's average launch time, measured with HIP events on the
launch stream, against 8 TB/s HBM.
cpu_baseline: the oracle (the CPU restatement of the reference's FP64
traversal) on a bounded sample of the same rays, on all host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np

HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY = 48
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_closest_trefoil_traffic.json")


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, rays_sets, budget_s=12.0):
    """Oracle FP64 closest hit on a bounded sample, all host cores."""
    import oracle
    o = oracle.OracleScene(scene.to_desc())
    threads = os.cpu_count() or 1
    # Calibrate on 2^16 evenly spaced rays per set, then size the sample to ~budget_s.
    t0 = time.perf_counter()
    for r in rays_sets:
        o.trace_closest(r[:: len(r) >> 16], fp32=False, threads=threads)
    rate = 2 * (1 << 16) / (time.perf_counter() - t0)
    per_set = int(min(len(rays_sets[0]), max(1 << 16, rate * budget_s / 2)))
    stride = max(1, len(rays_sets[0]) // per_set)
    # When both full sets take less than the budget, trace them several times.
    passes = max(1, int(rate * budget_s / (2 * len(rays_sets[0])))) if stride == 1 else 1
    done, elapsed = 0, 0.0
    for _ in range(passes):
        for r in rays_sets:
            sample = r[::stride]
            t0 = time.perf_counter()
            o.trace_closest(sample, fp32=False, threads=threads)
            elapsed += time.perf_counter() - t0
            done += len(sample)
    return {"value": round(done / elapsed / 1e6, 3), "unit": "Mrays/s", "cores": threads, "cpu": cpu_model(),
            "kind": "port",
            "sample": "every %d-th ray of both 2^24 sets, %d pass(es) (%d rays), FP64 oracle traversal, %.1f s"
                      % (stride, passes, done, elapsed)}


def path_tracing_leg(ctx, cpu=True):
    """Sphere room, 320x180 at 16 spp, max depth 8, roulette from bounce 3."""
    from albedo_amd import workloads as W
    from albedo_amd.scene import CameraSettings
    s = W.room_scene()
    g = ctx.scene(s)
    c = W.ROOM_CAMERA
    cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, 320 / 180)
    g.render(cam, 320, 180, 4, 8, 3, 0)  # warm-up (graph capture, allocation)
    t0 = time.perf_counter()
    reps = 20  # about 35 ms: enough renders that host jitter averages out
    st = None
    for r in range(reps):
        _, st = g.render(cam, 320, 180, 4, 8, 3, r, stats=True)
    wall = (time.perf_counter() - t0) / reps
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    out = {"workload": "sphere room 320x180, 16 spp, max depth 8, roulette from 3",
           "mpixel_spp_s": round(320 * 180 * 16 / wall / 1e6, 2), "mrays_s": round(rays / wall / 1e6, 2),
           "ms_per_image": round(wall * 1e3, 2), "rays_per_image": int(rays), "gpu_ms": round(st["gpu_ms"], 2)}
    if cpu:
        import oracle
        o = oracle.OracleScene(s.to_desc())
        t0 = time.perf_counter()
        _, counts = o.render(cam.to_abi(), W.render_params(320, 180, 4, 8, 3, 0), sample_mode=0)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"mpixel_spp_s": round(320 * 180 * 16 / dt / 1e6, 3),
                               "mrays_s": round(float(counts.sum()) / dt / 1e6, 3), "cores": os.cpu_count(),
                               "cpu": cpu_model(),
                               "kind": "port", "sample": "the full image, FP64 oracle estimator, reference draw order"}
    return out


def full_scene_leg(ctx, cpu=True):
    """The full-feature scene (the room with the checker floor, three smooth trefoil
    instances in gold, glass and copper, a Perlin marble sphere) at 1920x1080,
    256 spp, max depth 16, roulette from bounce 3: one render, after a small one
    that captures the bounce graph."""
    from albedo_amd import workloads as W
    from albedo_amd.scene import CameraSettings
    s = W.full_scene()
    g = ctx.scene(s)
    c = W.ROOM_CAMERA
    w, h, spa, depth = 1920, 1080, 16, 16
    cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, w / h)
    g.render(cam, 192, 108, 2, depth, 3, 0)
    t0 = time.perf_counter()
    _, st = g.render(cam, w, h, spa, depth, 3, 1, stats=True)
    wall = time.perf_counter() - t0
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    samples = w * h * spa * spa
    out = {"workload": "full scene 1920x1080, 256 spp, max depth 16, roulette from 3",
           "mpixel_spp_s": round(samples / wall / 1e6, 2), "mrays_s": round(rays / wall / 1e6, 2),
           "s_per_image": round(wall, 3), "rays_per_image": int(rays), "gpu_ms": round(st["gpu_ms"], 1),
           # The reference's README (README.md:74-75): meshes.scene, three trefoil instances on a
           # checker floor, 1920x1080 at 256 spp in about 9 min on an 8-thread CPU, with its
           # estimator's default max depth of 64 (this workload stops at 16).
           "reference_published": {"mpixel_spp_s": 0.98, "s_per_image": 540.0,
                                   "note": "meshes.scene, 8-thread CPU, max depth 64"}}
    g.release()
    print("full: rendered in %.2f s" % wall, file=sys.stderr, flush=True)
    if cpu:
        import oracle
        o = oracle.OracleScene(s.to_desc())
        cw, ch, cspa = 640, 360, 4
        ccam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, cw / ch)
        t0 = time.perf_counter()
        _, counts = o.render(ccam.to_abi(), W.render_params(cw, ch, cspa, depth, 3, 1), sample_mode=0)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"mpixel_spp_s": round(cw * ch * cspa * cspa / dt / 1e6, 3),
                               "mrays_s": round(float(counts.sum()) / dt / 1e6, 3), "cores": os.cpu_count(),
                               "cpu": cpu_model(),
                               "kind": "port",
                               "sample": "the same scene at 640x360, 16 spp, FP64 oracle estimator, reference draw order"}
    return out


def batch_leg(ctx, n_jobs=16, full=False, forms=("plain", "batched"), tune=True):
    """Mixed jobs (room / full-scene variants with 5..200 random spheres): one at a
    time vs one batched launch, each with its autotuned configuration (or, with
    tune=False, the context's default one). Default: 16 jobs, 256x144..960x540,
    4..64 spp; full: the 64-job mix of resolutions 256x144..1920x1080 and 4..256
    spp. `forms` picks which of the two forms run (for profiling one alone)."""
    import albedo_amd
    from albedo_amd import workloads as W
    from albedo_amd.scene import CameraSettings
    knot = W.trefoil_mesh()
    if full:
        mix = W.job_mix(n_jobs, knot=knot)
    else:
        mix = W.job_mix(n_jobs, resolutions=W.RESOLUTIONS[:3], spas=W.SPAS[:3], knot=knot)
    c = W.ROOM_CAMERA
    jobs = []
    for scene, (w, h), spa, depth, seed in mix:
        jobs.append((ctx.scene(scene), CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, w / h), w, h, spa,
                     depth, 3, seed))
    base = ctx.launch_config()
    default = {"block_size": base.block_size, "paths_in_flight": base.paths_in_flight,
               "grid_blocks": base.grid_blocks, "use_graph": base.use_graph}
    samples = sum(j[2] * j[3] * j[4] ** 2 for j in jobs)
    sizes = "256x144..1920x1080, 4..256 spp" if full else "256x144..960x540, 4..64 spp"
    out = {"workload": "%d mixed jobs, %s, 5..200 random spheres" % (n_jobs, sizes), "samples": samples}
    # Each form runs with the launch configuration the autotuner picks for it.
    if "plain" in forms:
        cfg_plain = ctx.autotune(jobs, batched=False)[0] if tune else default
        print("batch: tuned the plain form", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        rays_plain = 0
        plain_imgs = []
        for j in jobs:
            img, st = j[0].render(*j[1:], stats=True)
            plain_imgs.append(img)
            rays_plain += st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
        plain = time.perf_counter() - t0
        print("batch: %d jobs one at a time in %.2f s" % (len(jobs), plain), file=sys.stderr, flush=True)
        ctx.set_launch_config(base.block_size, base.paths_in_flight, base.grid_blocks, base.use_graph)
        out.update({"plain_s": round(plain, 3), "plain_mrays_s": round(rays_plain / plain / 1e6, 1),
                    "tuned_plain": cfg_plain})
    if "batched" in forms:
        cfg_batch = ctx.autotune(jobs, batched=True)[0] if tune else default
        print("batch: tuned the batched form", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        imgs, st = albedo_amd.render_batch(ctx, jobs)
        batched = time.perf_counter() - t0
        print("batch: one batched launch in %.2f s" % batched, file=sys.stderr, flush=True)
        ctx.set_launch_config(base.block_size, base.paths_in_flight, base.grid_blocks, base.use_graph)
        rays_b = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
        out.update({"batched_s": round(batched, 3), "batched_mrays_s": round(rays_b / batched / 1e6, 1),
                    "batched_mpixel_spp_s": round(samples / batched / 1e6, 1), "tuned_batched": cfg_batch})
    if "plain_s" in out and "batched_s" in out:
        out["speedup"] = round(out["plain_s"] / out["batched_s"], 3)
        same = sum(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(plain_imgs, imgs))
        out["results_equal_alone"] = "%d/%d" % (same, len(jobs))
    return out


def build_leg(ctx):
    """BVH build of the serving scene (1,310,720-triangle displaced icosphere x 64
    instances): scene creation with the host SAH BLAS, the device linear BVH
    (AB_BUILD_DEVICE_BLAS) and the host SAH trees collapsed four wide
    (AB_BUILD_WIDE_BVH), and closest-hit throughput of each on 2^22 rays into the
    instance grid (hits are identical; the trees differ)."""
    import torch
    from albedo_amd import workloads as W
    s = W.icosphere_scene()
    out = {"workload": "icosphere 1,310,720 triangles x 64 instances; 2^22 rays from the serving camera"}
    rng = np.random.default_rng(21)
    n = 1 << 22
    org = np.tile(np.array(W.ICOSPHERE_CAMERA.look_from, np.float32), (n, 1))
    tgt = np.stack([rng.uniform(-6, 6, n), rng.uniform(0.0, 1.5, n), rng.uniform(-6, 6, n)], 1)
    d = (tgt - org).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    d_rays = torch.from_numpy(rays).cuda()
    d_hits = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    digests = {}
    for name, device, wide in (("host_sah", False, False), ("device_lbvh", True, False), ("host_sah_wide", False, True)):
        t0 = time.perf_counter()
        g = ctx.scene(s, device_blas=device, wide_bvh=wide)
        build_s = time.perf_counter() - t0
        g.trace_closest_device(d_rays.data_ptr(), d_hits.data_ptr(), n, stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ms = []
        for _ in range(3):
            e0.record(stream)
            g.trace_closest_device(d_rays.data_ptr(), d_hits.data_ptr(), n, stream.cuda_stream)
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        digests[name] = hash(d_hits.cpu().numpy().tobytes())
        info = g.info()
        out[name] = {"scene_create_ms": round(build_s * 1e3, 1), "blas_nodes": info["blas_nodes"],
                     "stack_depth": info["stack_depth"], "mrays_s": round(n / (min(ms) * 1e-3) / 1e6, 1)}
        g.release()
        print("build: %s done" % name, file=sys.stderr, flush=True)
    out["hits_identical"] = digests["host_sah"] == digests["device_lbvh"] == digests["host_sah_wide"]
    return out


def _fifo_latencies(arrivals, service):
    """Latencies of a one-job-at-a-time server taking jobs in arrival order."""
    free, out = 0.0, []
    for a, s in zip(arrivals, service):
        free = max(free, a) + s
        out.append(free - a)
    return np.array(out)


def serve_leg(ctx, n_jobs=24, loads=(0.5, 1.0, 2.0), budget=1 << 30, runs=None):
    """Serving over a library of 8 HBM-resident scenes (the 1,310,720-triangle
    displaced icosphere x 64 instances on a checker floor, three sphere-room and
    four full-scene variants with 5..200 random spheres). Jobs: random scene,
    resolution 256x144..1920x1080, 4..256 spp. Poisson arrivals at rates swept as
    multiples of the plain path's measured capacity; admission by a 1 GiB HBM
    budget. The plain path (each job alone, one at a time, in arrival order) is
    timed on the same jobs and replayed against the same arrival times. `runs`
    (a list of (load, policy name)) replaces the default sweep."""
    import hashlib
    import albedo_amd
    from albedo_amd import _abi
    from albedo_amd import workloads as W
    from albedo_amd.scene import CameraSettings
    t0 = time.perf_counter()
    library = W.serving_library()
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    gscenes = [ctx.scene(s) for _, s, _, _ in library]
    t_load = time.perf_counter() - t0
    print("serve: library generated in %.1f s, built and resident in %.1f s" % (t_gen, t_load), file=sys.stderr,
          flush=True)

    def job_args(k, w, h, spa, seed):
        _, _, c, depth = library[k]
        cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), c.vertical_fov_degrees, w / h)
        return (gscenes[k], cam, w, h, spa, depth, 3, seed)

    jobs = [job_args(*j) for j in W.serving_jobs(n_jobs, len(library))]
    for k, g in enumerate(gscenes):  # first use of each scene's launch shapes
        g.render(*job_args(k, 64, 36, 1, 0)[1:])
    service, digests, rays_plain = [], [], 0
    for j in jobs:
        t1 = time.perf_counter()
        img, st = j[0].render(*j[1:], stats=True)
        service.append(time.perf_counter() - t1)
        rays_plain += st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
        digests.append(hashlib.sha1(img.tobytes()).hexdigest())
    capacity = n_jobs / sum(service)
    samples = sum(j[2] * j[3] * j[4] ** 2 for j in jobs)
    print("serve: plain path %.2f s for %d jobs" % (sum(service), n_jobs), file=sys.stderr, flush=True)
    sweep = []
    names = {_abi.SCHED_FIFO: "fifo", _abi.SCHED_SHARE: "share", _abi.SCHED_SMALL_FIRST: "small_first"}
    if runs is None:
        runs = [(load, _abi.SCHED_SMALL_FIRST) for load in loads]
        runs += [(1.0, p) for p in (_abi.SCHED_FIFO, _abi.SCHED_SHARE)]
    else:
        by_name = {v: k for k, v in names.items()}
        runs = [(load, by_name[p]) for load, p in runs]
    for load, policy in runs:
        rate = load * capacity
        arrivals = W.poisson_arrivals(n_jobs, rate)
        with albedo_amd.Server(ctx, budget, policy=policy) as srv:
            t0 = time.perf_counter()
            ids = []
            for j, at in zip(jobs, arrivals):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(*j))
            same = 0
            for i, d in zip(ids, digests):
                _, _, img = srv.wait(i)
                same += hashlib.sha1(img.tobytes()).hexdigest() == d
            wall = time.perf_counter() - t0
            st = srv.stats()
        plain_lat = _fifo_latencies(arrivals, service) * 1e3
        plain_wall = float(np.max(arrivals + plain_lat / 1e3))  # last result, from time 0 as `wall`
        sweep.append({"rate_jobs_s": round(rate, 3), "load": load, "policy": names[policy],
                      "p50_latency_ms": round(st["p50_latency_ms"], 1),
                      "p99_latency_ms": round(st["p99_latency_ms"], 1),
                      "mrays_s": round(st["rays"] / wall / 1e6, 1), "wall_s": round(wall, 2),
                      "plain_p50_latency_ms": round(float(np.percentile(plain_lat, 50)), 1),
                      "plain_p99_latency_ms": round(float(np.percentile(plain_lat, 99)), 1),
                      "plain_mrays_s": round(rays_plain / plain_wall / 1e6, 1),
                      "results_equal_alone": "%d/%d" % (same, n_jobs)})
        print("serve: load %.2f %s done in %.1f s" % (load, names[policy], wall), file=sys.stderr, flush=True)
    return {"workload": "%d jobs over 8 resident scenes (1,310,720-triangle displaced icosphere x 64 instances, "
                        "room and full-scene variants), 256x144..1920x1080, 4..256 spp, Poisson arrivals, "
                        "1 GiB HBM budget" % n_jobs,
            "scenes": [name for name, *_ in library], "library_build_s": round(t_gen + t_load, 1),
            "samples": samples, "plain_capacity_jobs_s": round(capacity, 3),
            "plain_mrays_s_busy": round(rays_plain / sum(service) / 1e6, 1), "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--trace-only", action="store_true", help="closest-hit benchmark only")
    ap.add_argument("--legs", default="pt,full,batch64,serve,build",
                    help="legs after the closest-hit benchmark, comma separated: pt (sphere room), full "
                         "(full-feature scene), batch64 (the 64-job mix up to 1920x1080 at 256 spp), batch "
                         "(a 16-job mix up to 960x540 at 64 spp), serve, build")
    ap.add_argument("--batch-forms", default="plain,batched",
                    help="forms the batch legs run: plain (one job at a time), batched, or both")
    ap.add_argument("--fixed-config", action="store_true",
                    help="batch legs: the context's default launch configuration instead of the autotuner's")
    ap.add_argument("--serve-runs", default=None,
                    help="serve leg: comma-separated load:policy runs (policy fifo, share or small_first) "
                         "instead of the default sweep")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy, the reference for counter-based traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this path has no multi-GPU form (each GPU would trace its own "
              "share of rays against a replicated scene); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import albedo_amd
    from albedo_amd import workloads as W

    scene, mesh = W.trefoil_scene()
    sets = {"coherent": W.coherent_rays(mesh, 4096), "incoherent": W.incoherent_rays(mesh, 1 << 24)}
    n_per = 1 << 24
    ctx = albedo_amd.Context(0)
    g = ctx.scene(scene)
    dev = torch.device("cuda:0")
    d_rays = {k: torch.from_numpy(v).to(dev) for k, v in sets.items()}
    d_hits = {k: torch.empty((n_per, 4), dtype=torch.int32, device=dev) for k in sets}
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step(events=None):
        for k in sets:
            if events is not None:
                events[k][0].record(stream)
            g.trace_closest_device(d_rays[k].data_ptr(), d_hits[k].data_ptr(), n_per, sp)
            if events is not None:
                events[k][1].record(stream)

    if args.copy_probe:
        # A copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch calibrate
        # the counters (256 MiB read + 256 MiB written).
        src = torch.empty(1 << 26, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    per_set = {k: [] for k in sets}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for k in sets}
        step(ev)
        for k in sets:
            per_set[k].append(ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in per_set.items()}
    rays = 2 * n_per * args.steps
    value = rays / wall / 1e6
    avg_launch_ms = float(np.mean(list(kernel_ms.values())))
    achieved = BYTES_PER_RAY * n_per / (avg_launch_ms * 1e-3) / 1e9
    # HBM bytes per launch need hardware counters (rocprofv3 --pmc FETCH_SIZE,
    # WRITE_SIZE over this same command), which a run cannot read of itself:
    # the figure comes from the committed counter summary and says so.
    traffic, traffic_source = None, None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        traffic_source = "profiles/%s (%s)" % (os.path.basename(TRAFFIC_FILE), tj.get("command", "rocprofv3 --pmc"))
    out = {
        "metric": "Mrays/s + Mpixel·spp/s on 1 MI355X; p50/p99 job latency, served vs plain",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "closest-hit microbenchmark: trefoil tube 512x32 segments (32768 FP32 triangles), "
                               "2^24 coherent + 2^24 incoherent rays per step",
                   "rays_per_step": 2 * n_per, "parallelism": "single GPU"},
        "per_set_mrays_s": {k: round(n_per / (v * 1e-3) / 1e6, 1) for k, v in kernel_ms.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": "kQuery (closest hit)", "avg_launch_ms": round(avg_launch_ms, 4)},
    }
    hits = d_hits["incoherent"][:4096].cpu().numpy().view(np.uint32)
    if not np.any(hits[:, 1] != 0xFFFFFFFF):
        raise RuntimeError("benchmark produced no hits")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, [sets["coherent"], sets["incoherent"]])
    del d_rays, d_hits
    torch.cuda.empty_cache()
    legs = set() if args.trace_only else set(args.legs.split(","))
    if "pt" in legs:
        out["path_tracing"] = path_tracing_leg(ctx, cpu=not args.no_cpu)
    if "full" in legs:
        out["full_scene"] = full_scene_leg(ctx, cpu=not args.no_cpu)
    forms = tuple(args.batch_forms.split(","))
    if "batch" in legs:
        out["batched"] = batch_leg(ctx, forms=forms, tune=not args.fixed_config)
    if "batch64" in legs:
        out["batched_64"] = batch_leg(ctx, 64, full=True, forms=forms, tune=not args.fixed_config)
    if "serve" in legs:
        runs = None
        if args.serve_runs:
            runs = [(float(r.split(":")[0]), r.split(":")[1]) for r in args.serve_runs.split(",")]
        out["served"] = serve_leg(ctx, runs=runs)
    if "build" in legs:
        out["bvh_build"] = build_leg(ctx)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
