"""GPU tests of the serving loop (ab_server_*, ab_job_*): results equal the
plain renderer's bit for bit, the HBM budget admits and rejects, cancellation,
progressive snapshots, latency statistics."""
import time

import numpy as np
import pytest

import albedo_amd
from albedo_amd import AlbedoError, Server, _abi
from albedo_amd import workloads as W
from albedo_amd.scene import CameraSettings

pytestmark = pytest.mark.gpu


def cam(aspect, fov=38.0):
    c = W.ROOM_CAMERA
    return CameraSettings(c.look_from, c.look_at, (0, 1, 0), fov, aspect)


@pytest.fixture(scope="module")
def scenes(ctx):
    return ctx.scene(W.room_scene()), ctx.scene(W.full_scene())


def test_served_jobs_equal_plain_renders(ctx, scenes):
    room, full = scenes
    jobs = [(room, cam(40 / 24), 40, 24, 2, 8, 3, 1), (full, cam(64 / 36), 64, 36, 3, 16, 3, 2),
            (room, cam(1.0, 50.0), 17, 17, 4, 4, 0, 9), (full, cam(2.0), 32, 16, 2, 8, 3, 4)]
    with Server(ctx, 1 << 30) as srv:
        ids = [srv.submit(*j) for j in jobs]
        results = [srv.wait(i) for i in ids]
        st = srv.stats()
    for (g, c, w, h, n, d, rr, seed), (code, status, img) in zip(jobs, results):
        assert code == _abi.AB_OK and status["state"] == Server.DONE
        assert status["samples_done"] == w * h * n * n
        single = g.render(c, w, h, n, d, rr, seed)
        assert np.array_equal(single.view(np.uint32), img.view(np.uint32))
    assert st["jobs_done"] == len(jobs)
    assert st["p99_latency_ms"] >= st["p50_latency_ms"] > 0
    assert st["rays"] > 0


def test_budget_rejects_and_serialises(ctx, scenes):
    room, _ = scenes
    one_job = Server.footprint(room, cam(1.0), 64, 64, 4, 8)
    # accumulators (32 B per pixel), counter, mean image (12 B per pixel)
    assert one_job >= 64 * 64 * 44 + 8
    assert Server.footprint(room, cam(1.0), 128, 128, 1, 4) > one_job
    with Server(ctx, one_job) as srv:
        with pytest.raises(AlbedoError, match="AB_REJECTED_BUDGET"):
            srv.submit(room, cam(1.0), 128, 128, 1, 4)
        a = srv.submit(room, cam(1.0), 64, 64, 4, 8)
        b = srv.submit(room, cam(1.0), 64, 64, 4, 8, seed=1)
        _, sa, _ = srv.wait(a)
        _, sb, _ = srv.wait(b)
        # Only one fits the budget: the second starts after the first finished.
        assert sb["start_ms"] >= sa["finish_ms"]
        assert srv.stats()["jobs_rejected"] == 1


def test_cancel_and_progressive_snapshots(ctx, scenes):
    room, full = scenes
    with Server(ctx, 1 << 31) as srv:
        big = srv.submit(full, cam(16 / 9), 640, 360, 16, 16)
        small = srv.submit(room, cam(1.0), 32, 32, 2, 8)
        seen = []
        for _ in range(200):
            code, st, img = srv.poll(big)
            seen.append(st["samples_done"])
            if st["samples_done"] > 0:
                break
            time.sleep(0.01)
        srv.cancel(big)
        code, st, _ = srv.wait(big)
        assert code == _abi.AB_CANCELLED and st["state"] == Server.CANCELLED
        code, st, img = srv.wait(small)
        assert code == _abi.AB_OK
        assert np.array_equal(img.view(np.uint32), room.render(cam(1.0), 32, 32, 2, 8).view(np.uint32))
        assert seen == sorted(seen)
        srv.release(big)
        with pytest.raises(AlbedoError, match="AB_NOT_FOUND"):
            srv.poll(big)


def test_cancelled_jobs_release_their_memory_under_load(ctx, scenes):
    # A long job keeps the server busy while two others are admitted and
    # cancelled: their buffers go at once (charged bytes and device free memory
    # both drop while the long job still runs), and two new jobs fill the budget
    # they left and render as alone.
    import torch
    room, _ = scenes
    size = (1920, 1080)
    fp = Server.footprint(room, cam(16 / 9), *size, 1, 8)
    with Server(ctx, 3 * fp) as srv:
        long_job = srv.submit(room, cam(16 / 9), *size, 16, 8)
        a = srv.submit(room, cam(16 / 9), *size, 16, 8)
        b = srv.submit(room, cam(16 / 9), *size, 16, 8)
        for _ in range(500):
            if all(srv.poll(j, image=False)[1]["state"] == Server.RUNNING for j in (long_job, a, b)):
                break
            time.sleep(0.002)
        assert srv.stats()["hbm_charged"] == 3 * fp
        torch.cuda.synchronize()
        free_before = torch.cuda.mem_get_info()[0]
        srv.cancel(a)
        srv.cancel(b)
        assert srv.wait(a)[0] == _abi.AB_CANCELLED and srv.wait(b)[0] == _abi.AB_CANCELLED
        assert srv.poll(long_job, image=False)[1]["state"] == Server.RUNNING
        assert srv.stats()["hbm_charged"] == fp
        assert torch.cuda.mem_get_info()[0] - free_before >= 1.5 * fp
        c = srv.submit(room, cam(1.0), 64, 64, 2, 8)
        d = srv.submit(room, cam(1.0), 48, 48, 2, 8, seed=3)
        for j, w, seed in ((c, 64, 0), (d, 48, 3)):
            code, _, img = srv.wait(j)
            assert code == _abi.AB_OK
            alone = room.render(cam(1.0), w, w, 2, 8, seed=seed)
            assert np.array_equal(img.view(np.uint32), alone.view(np.uint32))
        assert srv.wait(long_job)[0] == _abi.AB_OK
        assert srv.stats()["hbm_charged"] == 0


def test_progressive_snapshot_converges_to_final(ctx, scenes):
    room, _ = scenes
    with Server(ctx, 1 << 30) as srv:
        j = srv.submit(room, cam(16 / 9), 160, 90, 16, 8)
        partial = None
        for _ in range(500):
            code, st, img = srv.poll(j)
            if code == _abi.AB_PENDING and 0 < st["samples_done"] < st["samples_total"]:
                partial = img
            if code == _abi.AB_OK:
                break
            time.sleep(0.002)
        code, st, final = srv.wait(j)
        assert code == _abi.AB_OK
        if partial is not None:
            assert np.all(np.isfinite(partial))
            rel = abs(float(partial.mean()) - float(final.mean())) / float(final.mean())
            assert rel < 0.5


def test_mix_with_million_triangle_scene_equals_alone(ctx, scenes, icosphere):
    # Jobs of the sphere room (20-entry stack class) are in flight when jobs of the
    # 1.3M-triangle instanced scene (64-entry class) arrive and upgrade the shared
    # launch; each result still equals its render alone.
    room, full = scenes
    _, ico, _ = icosphere
    c = W.ICOSPHERE_CAMERA
    ico_cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), c.vertical_fov_degrees, 96 / 54)
    jobs = [(room, cam(160 / 90), 160, 90, 4, 8, 3, 21), (ico, ico_cam, 96, 54, 3, 8, 3, 22),
            (full, cam(64 / 36), 64, 36, 2, 16, 3, 23), (ico, ico_cam, 96, 54, 2, 8, 3, 24),
            (room, cam(1.0), 24, 24, 2, 8, 3, 25)]
    with Server(ctx, 1 << 30) as srv:
        ids = [srv.submit(*j) for j in jobs]
        results = [srv.wait(i) for i in ids]
    for (g, cm, w, h, n, d, rr, seed), (code, status, img) in zip(jobs, results):
        assert code == _abi.AB_OK and status["samples_done"] == w * h * n * n
        alone = g.render(cm, w, h, n, d, rr, seed)
        assert np.array_equal(alone.view(np.uint32), img.view(np.uint32))


def test_one_mesh_scene_in_flight_then_a_cancel_and_other_scenes(ctx, scenes, icosphere):
    # While only jobs of the icosphere scene are in flight the bounces trace over
    # that one scene; a cancelled job keeps every scene of the table traced until
    # its paths drain, and a room job brings the split tracing back. Every
    # finished job equals its render alone.
    room, _ = scenes
    _, ico, _ = icosphere
    c = W.ICOSPHERE_CAMERA
    ico_cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), c.vertical_fov_degrees, 96 / 54)
    first = [(ico, ico_cam, 96, 54, 3, 8, 3, 31), (ico, ico_cam, 64, 36, 2, 8, 3, 32)]
    later = [(room, cam(1.0), 24, 24, 2, 8, 3, 34), (ico, ico_cam, 48, 27, 2, 8, 3, 35)]
    with Server(ctx, 1 << 31) as srv:
        results = [srv.wait(i) for i in [srv.submit(*j) for j in first]]
        big = srv.submit(ico, ico_cam, 640, 360, 16, 8, 3, 33)
        for _ in range(200):
            if srv.poll(big)[1]["samples_done"] > 0:
                break
            time.sleep(0.005)
        srv.cancel(big)
        ids = [srv.submit(*j) for j in later]
        code, st, _ = srv.wait(big)
        assert code == _abi.AB_CANCELLED and st["state"] == Server.CANCELLED
        results += [srv.wait(i) for i in ids]
    for (g, cm, w, h, n, d, rr, seed), (code, status, img) in zip(first + later, results):
        assert code == _abi.AB_OK and status["samples_done"] == w * h * n * n
        alone = g.render(cm, w, h, n, d, rr, seed)
        assert np.array_equal(alone.view(np.uint32), img.view(np.uint32))


@pytest.mark.parametrize("policy", [_abi.SCHED_FIFO, _abi.SCHED_SHARE, _abi.SCHED_SMALL_FIRST])
def test_policies_keep_images_and_order_jobs(ctx, scenes, policy):
    room, full = scenes
    big = (full, cam(16 / 9), 640, 360, 16, 16, 3, 31)    # 59M samples
    small = (room, cam(1.0), 32, 32, 2, 8, 3, 32)          # 4096 samples
    with Server(ctx, 1 << 30, policy=policy) as srv:
        a = srv.submit(*big)
        b = srv.submit(*small)
        _, sa, ia = srv.wait(a)
        _, sb, ib = srv.wait(b)
    assert np.array_equal(ia.view(np.uint32), full.render(*big[1:]).view(np.uint32))
    assert np.array_equal(ib.view(np.uint32), room.render(*small[1:]).view(np.uint32))
    if policy != _abi.SCHED_FIFO:
        # The small job does not wait for the large one's samples.
        assert sb["finish_ms"] < sa["finish_ms"]


def test_unknown_policy_is_rejected(ctx):
    with Server(ctx, 1 << 30) as srv:
        with pytest.raises(AlbedoError, match="unknown policy"):
            srv.set_policy(7)


def test_wire_format_sums_are_progressive_and_exact(ctx, scenes):
    room, _ = scenes
    job = (room, cam(16 / 9), 160, 90, 8, 8, 3, 41)
    with Server(ctx, 1 << 30) as srv:
        j = srv.submit(*job, keep_sums=True)
        k = srv.submit(*job)  # the same job without kept sums
        counts = []
        for _ in range(400):
            code, st, sums = srv.poll_sums(j)
            counts.append(sums[..., 3].astype(np.int64))
            if code == _abi.AB_OK:
                break
            time.sleep(0.002)
        code, st, img = srv.wait(j)
        assert code == _abi.AB_OK
        _, _, final = srv.poll_sums(j)
        srv.wait(k)
        with pytest.raises(AlbedoError, match="AB_JOB_KEEP_SUMS"):
            srv.poll_sums(k)
    # Per-pixel counts never go down between snapshots and end at spp everywhere.
    for a, b in zip(counts, counts[1:]):
        assert np.all(b >= a)
    assert np.all(final[..., 3] == 64)
    # The final sums resolve on the host to the served image, bit for bit.
    assert np.array_equal(albedo_amd.sums_resolve(final).view(np.uint32), img.view(np.uint32))
    assert np.array_equal(img.view(np.uint32), room.render(*job[1:]).view(np.uint32))
