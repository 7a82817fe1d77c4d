"""The serving scheduler's host logic on the CPU, with stand-in slots whose
phases take a scripted number of polls and a scripted number of bytes:
continuous batching, first-come-first-served admission, the HBM budget never
exceeded, preemption when shading memory runs out, rejection of requests that
cannot fit alone."""
import itertools

import numpy as np
import pytest

from nsvf_amd.serve import HbmBudget, RequestTooLarge, Scheduler, Slot, latency_summary


class FakeSlot(Slot):
    """Geometry takes `g` polls and R*gb bytes; shading `s` polls and
    shade[rid] (default R*sb) bytes."""

    def __init__(self, g=2, s=3, gb=10, sb=20, shade=None):
        self.g, self.s, self.gb, self.sb = g, s, gb, sb
        self.shade = shade or {}
        self.req = None
        self.phase = None
        self.t = 0

    def geometry_bytes(self, req):
        return req.rays * self.gb

    def start_geometry(self, req):
        self.req, self.phase, self.t = req, "geometry", 0

    def geometry_ready(self):
        self.t += 1
        return self.t > self.g

    def shading_bytes(self):
        return self.shade.get(self.req.rid, self.req.rays * self.sb)

    def start_shading(self):
        self.phase, self.t = "shading", 0

    def shading_done(self):
        self.t += 1
        return self.t > self.s

    def result(self):
        return ("image", self.req.rid)

    def reset(self):
        self.req, self.phase, self.t = None, None, 0


def rays(n):
    return np.zeros((n, 3), np.float32)


def make(n_slots, budget, **kw):
    clock = itertools.count()
    return Scheduler([FakeSlot(**kw) for _ in range(n_slots)], HbmBudget(budget),
                     clock=lambda: next(clock))


def run(sched, max_polls=10_000):
    peak_in_flight = 0
    for _ in range(max_polls):
        if sched.idle():
            return peak_in_flight
        sched.poll()
        peak_in_flight = max(peak_in_flight, sched.in_flight)
        assert sched.budget.used <= sched.budget.total
    raise AssertionError("scheduler did not drain")


def test_all_complete_fifo_and_within_budget():
    sched = make(3, 10**9)
    reqs = [sched.submit("a", rays(100 + i), rays(100 + i)) for i in range(10)]
    peak = run(sched)
    assert peak == 3
    assert all(r.state == "done" and r.result == ("image", r.rid) for r in reqs)
    admitted = [r.admitted for r in reqs]
    assert admitted == sorted(admitted)
    assert sched.budget.used == 0


def test_continuous_batching_admits_as_slots_free():
    """A queued request starts when any slot frees, before the rest of the
    earlier batch has finished (not batch-at-a-time)."""
    sched = make(2, 10**9)
    slow = FakeSlot(g=1, s=30)
    fast = FakeSlot(g=1, s=1)
    sched.slots[0].slot, sched.slots[1].slot = slow, fast
    r0 = sched.submit("a", rays(10), rays(10))
    r1 = sched.submit("a", rays(10), rays(10))
    r2 = sched.submit("a", rays(10), rays(10))
    run(sched)
    assert r2.admitted < r0.finished, "r2 waited for the whole first batch"
    assert all(r.state == "done" for r in (r0, r1, r2))


def test_budget_serialises_requests_and_none_exceeds_it():
    # each request needs 1000 + 2000 bytes: two geometries fit, but only one
    # request can shade at a time (the other is preempted when it blocks)
    sched = make(4, 3500, gb=10, sb=20)
    reqs = [sched.submit("a", rays(100), rays(100)) for _ in range(5)]
    shading = 0
    while not sched.idle():
        sched.poll()
        shading = max(shading, sum(st.phase == "shading" for st in sched.slots))
        assert sched.budget.used <= 3500
    assert shading == 1
    assert sched.budget.peak <= 3500
    assert all(r.state == "done" for r in reqs)
    assert sum(r.preemptions for r in reqs) >= 1


def test_request_too_large_is_rejected_and_others_complete():
    sched = make(2, 5000, gb=10, sb=20)
    big = sched.submit("a", rays(1000), rays(1000))        # 10000 bytes of geometry
    small = [sched.submit("a", rays(50), rays(50)) for _ in range(3)]
    run(sched)
    assert big.state == "rejected" and "geometry needs" in big.error
    assert all(r.state == "done" for r in small)


def test_shading_too_large_alone_is_rejected():
    sched = make(1, 5000, gb=10, sb=20, shade={0: 10**6})
    r0 = sched.submit("a", rays(10), rays(10))
    r1 = sched.submit("a", rays(10), rays(10))
    run(sched)
    assert r0.state == "rejected" and "shading needs" in r0.error
    assert r1.state == "done"


def test_preemption_frees_geometry_memory_for_shading():
    """Two requests hold geometry memory; the first one's shading only fits
    if the second lets go: the second is preempted, re-queued first, and
    completes afterwards."""
    # geometry 1000 each; shading of r0 needs 2500: 1000 + 1000 + 2500 > 4000
    sched = make(2, 4000, gb=10, sb=5, shade={0: 2500})
    r0 = sched.submit("a", rays(100), rays(100))
    r1 = sched.submit("a", rays(100), rays(100))
    run(sched)
    assert r0.state == "done" and r1.state == "done"
    assert r1.preemptions >= 1 and r0.preemptions == 0
    assert sched.budget.used == 0


def test_budget_accounting_and_summary():
    b = HbmBudget(100)
    b.reserve(60)
    with pytest.raises(RequestTooLarge):
        b.reserve(50)
    b.release(60)
    assert b.used == 0 and b.peak == 60
    sched = make(2, 10**9)
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(4)]
    run(sched)
    s = latency_summary(reqs)
    assert s["done"] == 4 and s["rejected"] == 0 and s["p99_ms"] >= s["p50_ms"] > 0
