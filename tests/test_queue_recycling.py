"""A CPU model of the path tracer's queue bookkeeping with slot recycling.

The bounce kernels (albedo_amd/csrc/wavefront.hip: kGenerate, kPlanBins,
kScatterShade, kFinish, pendingGeneration) keep paths in `capacity` slots. While
the work left covers every path of a bounce, a terminated path keeps its queue
position, and the next bounce's first new samples take over its slot and
position (in termination-queue order); otherwise survivors are compacted and
freed slots go to a free stack, from which new samples are appended; before
paths are appended from a stack that a finish has pushed to, the stack is
rebuilt from the slots' free flags so it pops ascending slots. The model
follows those steps over random termination patterns, capacities and work that
arrives while the engine runs (as the server exposes it), and checks that no
live slot is ever handed out twice, that the queue names each live path once,
that occupied plus free slots always make up the capacity, and that every work
index is generated exactly once. It also checks what recycling and the
free-slot sort are for: until the work left no longer covers the queue,
position p holds slot p, and the paths of one generation that are appended
take ascending slots. The planted
mistake recycles without checking that the work left covers the terminated
paths, which leaves a position whose path never comes.
"""
import random

import pytest


def simulate(cap, arrivals, seed, check_work_left=True, sort_free=True, max_bounces=100000):
    """arrivals: work_total values, the first at the start, the next every third
    bounce. Returns (bounces, identity_held)."""
    rnd = random.Random(seed)
    freeq = [cap - 1 - i for i in range(cap)]  # kReset: the first generation takes slots 0, 1, 2, ...
    slot_free = [True] * cap
    c = dict(n_active=0, n_term=0, n_free=cap, n_recycle=0, work_next=0, work_total=arrivals[0], free_sorted=True)
    active, active_next, termq = [None] * cap, [None] * cap, [None] * cap
    live = {}  # slot -> work index
    generated = []
    later = list(arrivals[1:])
    identity = True
    recycled_so_far = True
    for bounce in range(1, max_bounces):
        if later and bounce % 3 == 0:
            c["work_total"] = later.pop(0)
        # pendingGeneration + kGenerate
        free_top = c["n_free"] + c["n_term"] - c["n_recycle"]
        rest = c["work_total"] - c["work_next"] - c["n_recycle"]
        if rest < 0:
            raise AssertionError("recycled positions left without work")
        extra = min(rest, free_top)
        n_new = c["n_recycle"] + extra
        if sort_free and extra > 0 and not c["free_sorted"]:  # kFreeCount + kFreeScatter
            free = [s for s in range(cap) if slot_free[s]]
            assert len(free) == free_top
            for r, s in enumerate(free):
                freeq[free_top - 1 - r] = s
        appended = []
        for i in range(n_new):
            if i < c["n_recycle"]:
                pos = termq[i]
                slot = active[pos]
            else:
                pos = c["n_active"] + (i - c["n_recycle"])
                slot = freeq[free_top - 1 - (i - c["n_recycle"])]
                slot_free[slot] = False
                appended.append(slot)
            assert slot not in live, "a live slot handed out again"
            live[slot] = c["work_next"] + i
            generated.append(c["work_next"] + i)
            active[pos] = slot
        assert appended == sorted(appended), "appended paths take ascending slots"
        # kPlanBins bookkeeping
        n = c["n_active"] + extra
        c.update(n_term=0, n_recycle=0, n_free=free_top - extra, n_active=n)
        if extra > 0:
            c["free_sorted"] = True
        c["work_next"] += n_new
        queue = active[:n]
        assert len(set(queue)) == n and all(s in live for s in queue)
        assert n + c["n_free"] == cap
        if recycled_so_far and queue != list(range(n)):
            identity = False
        # kShade + kScatterShade
        term = [rnd.random() < 0.3 for _ in range(n)]
        left = c["work_total"] - c["work_next"]
        recycle = left >= n if check_work_left else True
        recycled_so_far = recycled_so_far and recycle
        tpos = [i for i in range(n) if term[i]]
        if recycle:
            active_next[:n] = active[:n]
            c["n_active"] = n
        else:
            surv = [active[i] for i in range(n) if not term[i]]
            active_next[:len(surv)] = surv
            c["n_active"] = len(surv)
        termq[:len(tpos)] = tpos
        c["n_term"] = len(tpos)
        c["n_recycle"] = len(tpos) if recycle else 0
        # kFinish
        if c["n_term"] > c["n_recycle"]:
            c["free_sorted"] = False
        for k in range(c["n_term"]):
            slot = active[termq[k]]
            del live[slot]
            if k >= c["n_recycle"]:
                freeq[c["n_free"] + (k - c["n_recycle"])] = slot
                slot_free[slot] = True
        active, active_next = active_next, active
        if c["work_next"] == c["work_total"] and c["n_active"] == 0 and not later:
            assert sorted(generated) == list(range(c["work_total"])), "work generated once each"
            return bounce, identity
    raise AssertionError("did not drain")


@pytest.mark.parametrize("arrive", [False, True])
def test_bookkeeping_invariants(arrive):
    for seed in range(150):
        r = random.Random(seed)
        cap = r.choice([7, 64, 100, 256])
        if arrive:
            arrivals = sorted(r.sample(range(1, 3000), 5))
        else:
            arrivals = [r.randint(0, 2000)]
        simulate(cap, arrivals, seed)


def test_queue_position_holds_its_slot_while_work_lasts():
    # A render (all work known at the start): the queue is slots 0..n-1 in order
    # at every bounce until the work left no longer covers the queue.
    for seed in range(20):
        _, identity = simulate(128, [5000], seed)
        assert identity


def test_planted_recycling_without_the_work_check_fails():
    failures = 0
    for seed in range(30):
        try:
            simulate(64, [300], seed, check_work_left=False)
        except AssertionError:
            failures += 1
    assert failures > 0


def test_planted_append_without_the_free_sort_scatters_slots():
    # Work that arrives after the queue ran short: without the sort, paths
    # appended from the free stack take its pushes' order.
    failures = 0
    for seed in range(30):
        try:
            simulate(64, [100, 400, 900, 1500], seed, sort_free=False)
        except AssertionError as e:
            failures += "ascending" in str(e)
    assert failures > 0
