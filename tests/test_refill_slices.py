"""A CPU model of the lane-refill wave loop of the tracing kernels.

kTraceQueue (albedo_amd/csrc/wavefront.hip) and kQuery (trace.hip) run one wave
per slice of a queue: idle lanes take the slice's next rays once at least 16 of
the 64 lanes are free (or all are), each lane runs up to 16 traversal steps
between checks, and in split mode (a queue of several scenes) a lane skips a
ray that another kernel traces. The model follows that loop step for step on
random slices of 1-6 refills, random ownership and random traversal lengths,
and checks that every ray the kernel owns is traced exactly once and that the
loop ends. The planted mistake is the exit the kernel had before its fix --
leave as soon as no lane holds a ray -- which loses the rest of a slice when a
whole refill belongs to the other kernels; the model shows it does, and that it
was harmless where every ray is owned (kQuery).
"""
import numpy as np
import pytest

LANES, REFILL, STEPS = 64, 16, 16


def wave_loop(owned, steps_needed, fixed_exit=True, max_iters=100000):
    """Rays of the slice traced by one wave, in completion order."""
    end, nxt = len(owned), 0
    lane_id = [None] * LANES
    left = [0] * LANES
    traced = []
    for _ in range(max_iters):
        idle = [i is None for i in lane_id]
        free = sum(idle)
        if nxt < end and (free >= REFILL or free == LANES):
            rank = 0
            for lane in range(LANES):
                if not idle[lane]:
                    continue
                mine = nxt + rank
                rank += 1
                if mine < end and owned[mine]:
                    lane_id[lane] = mine
                    left[lane] = steps_needed[mine]
            nxt += free
        busy = any(i is not None for i in lane_id)
        if not busy and (nxt >= end or not fixed_exit):
            return traced
        for lane in range(LANES):
            if lane_id[lane] is None:
                continue
            left[lane] -= STEPS
            if left[lane] <= 0:
                traced.append(lane_id[lane])
                lane_id[lane] = None
    raise AssertionError("wave loop did not end")


def random_slice(rng, refills, owned_share, runs):
    n = LANES * refills - int(rng.integers(0, LANES))
    if runs:
        # Ownership in runs, as in a queue ordered by job: whole refills of
        # the other kernels' rays happen.
        owned = np.zeros(n, bool)
        pos = 0
        while pos < n:
            run = int(rng.integers(1, 3 * LANES))
            owned[pos:pos + run] = rng.random() < owned_share
            pos += run
    else:
        owned = rng.random(n) < owned_share
    steps = rng.geometric(1 / 60.0, n)
    return owned, steps


@pytest.mark.parametrize("runs", [False, True])
def test_every_owned_ray_is_traced_once(runs):
    rng = np.random.default_rng(7 if runs else 8)
    for _ in range(300):
        owned, steps = random_slice(rng, int(rng.integers(1, 7)), float(rng.choice([0.1, 0.5, 0.9, 1.0])), runs)
        traced = wave_loop(owned, steps)
        assert sorted(traced) == list(np.nonzero(owned)[0])


def test_planted_old_exit_loses_rays_only_in_split_mode():
    rng = np.random.default_rng(9)
    lost = 0
    for _ in range(300):
        owned, steps = random_slice(rng, int(rng.integers(2, 7)), 0.5, True)
        lost += len(np.nonzero(owned)[0]) - len(wave_loop(owned, steps, fixed_exit=False))
    assert lost > 0
    # Where the kernel owns every ray (the ray-query kernel), a refill with all
    # lanes free always hands out a ray, so the old exit was exact there.
    for _ in range(100):
        owned, steps = random_slice(rng, int(rng.integers(1, 7)), 1.0, False)
        assert len(wave_loop(owned, steps, fixed_exit=False)) == len(owned)


# ------------------------------------------------ scene-grouped lane refill
#
# kTraceGrouped (wavefront.hip) traces, in a queue of several scenes, the rays of
# the scenes with meshes, one scene per wave at a time: the queue keeps
# generation order, so a job's rays form runs. When every lane is idle the wave
# starts a run at its slice's next ray of such a scene (skipping rays of
# sphere-only scenes, which other kernels trace) and holds that scene in
# wave-uniform registers; later refills take only the run's following rays of the
# same scene and stop at the first other one.

def grouped_loop(keys, steps_needed, max_iters=200000):
    """keys[p]: the scene of position p, or None for rays of other kernels.
    Returns [(position, scene the wave held while tracing it)]."""
    end, nxt = len(keys), 0
    lane_id = [None] * LANES
    left = [0] * LANES
    cur = None
    traced = []
    for _ in range(max_iters):
        idle = [i is None for i in lane_id]
        free = sum(idle)
        if nxt < end and (free >= REFILL or free == LANES):
            cnt = min(free, end - nxt)
            window = keys[nxt:nxt + cnt]
            start = 0
            if free == LANES:
                owned = [k is not None for k in window]
                if not any(owned):
                    nxt += cnt
                    continue
                start = owned.index(True)
                cur = window[start]
            take = start
            while take < cnt and window[take] == cur:
                take += 1
            rank = 0
            for lane in range(LANES):
                if idle[lane]:
                    k = start + rank
                    rank += 1
                    if k < take:
                        lane_id[lane] = nxt + k
                        left[lane] = steps_needed[nxt + k]
            nxt += take
        if not any(i is not None for i in lane_id) and nxt >= end:
            return traced
        for lane in range(LANES):
            if lane_id[lane] is None:
                continue
            left[lane] -= STEPS
            if left[lane] <= 0:
                traced.append((lane_id[lane], cur))
                lane_id[lane] = None
    raise AssertionError("grouped wave loop did not end")


def random_keyed_slice(rng, refills, scenes, other_share):
    n = LANES * refills - int(rng.integers(0, LANES))
    keys, pos = [], 0
    while pos < n:
        run = int(rng.integers(1, 4 * LANES))
        k = None if rng.random() < other_share else int(rng.integers(0, scenes))
        keys += [k] * min(run, n - pos)
        pos += run
    return keys, rng.geometric(1 / 60.0, n)


def test_grouped_refill_traces_every_ray_once_over_its_own_scene():
    rng = np.random.default_rng(11)
    for _ in range(300):
        keys, steps = random_keyed_slice(rng, int(rng.integers(1, 7)), int(rng.integers(1, 5)),
                                         float(rng.choice([0.0, 0.3, 0.7])))
        traced = grouped_loop(keys, steps)
        assert sorted(p for p, _ in traced) == [p for p, k in enumerate(keys) if k is not None]
        # A lane only ever traced a ray of the scene its wave held.
        assert all(keys[p] == held for p, held in traced)


def test_grouped_refill_of_a_queue_of_one_other_kernel_ends():
    keys = [None] * (6 * LANES)
    assert grouped_loop(keys, [1] * len(keys)) == []
