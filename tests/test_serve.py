"""The serving scheduler's host logic on the CPU, with stand-in slots whose
phases take a scripted number of polls and a scripted number of bytes:
continuous batching, first-come-first-served admission, the HBM budget never
exceeded, preemption when shading memory runs out, rejection of requests that
cannot fit alone."""
import itertools

import numpy as np
import pytest

from nsvf_amd.serve import (HbmBudget, RequestTooLarge, Scheduler, Slot, SlotOutOfMemory,
                            latency_summary)


class FakeSlot(Slot):
    """Geometry takes `g` polls and R*gb bytes; shading `s` polls and
    shade[rid] (default R*sb) bytes.  With keep, the slot keeps its geometry
    and its shading allocation between jobs, each as large as it has been
    (as a GPU slot keeps its named workspace buffers), and a phase needs
    only what exceeds its kind."""

    def __init__(self, g=2, s=3, gb=10, sb=20, shade=None, keep=False):
        self.g, self.s, self.gb, self.sb = g, s, gb, sb
        self.shade = shade or {}
        self.keep = keep
        self.req = None
        self.phase = None
        self.t = 0
        self.kept = [0, 0]   # geometry, shading bytes kept between jobs
        self.live = [0, 0]   # of the job in flight

    def _grow(self, kind, n):
        return max(0, n - self.kept[kind]) if self.keep else n

    def geometry_bytes(self, job):
        return self._grow(0, job.rays * self.gb)

    def start_geometry(self, job):
        self.req, self.phase, self.t = job, "geometry", 0
        self.live = [job.rays * self.gb, 0]

    def geometry_ready(self):
        self.t += 1
        return self.t > self.g

    def _shade(self):
        return self.shade.get(self.req.rid, self.req.rays * self.sb)

    def shading_bytes(self):
        return self._grow(1, self._shade())

    def start_shading(self):
        self.phase, self.t = "shading", 0
        self.live[1] = self._shade()

    def shading_done(self):
        self.t += 1
        return self.t > self.s

    def result(self):
        return [("image", r.rid) for r in self.req.requests]

    def reset(self):
        if self.keep:
            self.kept = [max(k, l) for k, l in zip(self.kept, self.live)]
        self.req, self.phase, self.t, self.live = None, None, 0, [0, 0]

    def workspace_bytes(self):
        if self.keep:
            return sum(max(k, l) for k, l in zip(self.kept, self.live))
        return sum(self.live)

    def release_workspace(self):
        self.kept = [0, 0]


def rays(n):
    return np.zeros((n, 3), np.float32)


def make(n_slots, budget, max_batch=1, max_batch_rays=None, **kw):
    clock = itertools.count()
    return Scheduler([FakeSlot(**kw) for _ in range(n_slots)], HbmBudget(budget),
                     clock=lambda: next(clock), max_batch=max_batch, max_batch_rays=max_batch_rays)


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


def test_batches_group_requests_of_one_scene():
    """One slot, batches of up to 4: the oldest request and the later ones
    of its scene run together, each request getting its own result."""
    sched = make(1, 10**9, max_batch=4)
    scenes = ["a", "b", "a", "a", "b", "a", "a"]
    reqs = [sched.submit(sc, rays(10), rays(10)) for sc in scenes]
    run(sched)
    assert all(r.state == "done" and r.result == ("image", r.rid) for r in reqs)
    assert [r.batch for r in reqs] == [4, 2, 4, 4, 2, 4, 1]
    assert reqs[0].admitted == reqs[2].admitted == reqs[5].admitted
    assert reqs[1].admitted == reqs[4].admitted > reqs[0].admitted
    s = latency_summary(reqs)
    assert s["mean_batch"] == pytest.approx((4 * 4 + 2 * 2 + 1) / 7, abs=0.01)
    assert set(s["phases_ms"]) == {"queued", "geometry", "shading"}


def test_batches_stay_within_ray_cap_and_budget():
    sched = make(1, 10**9, max_batch=8, max_batch_rays=25)
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(5)]
    run(sched)
    assert [r.batch for r in reqs] == [2, 2, 2, 2, 1]
    # geometry of 3 requests (300 bytes) does not fit a 250-byte budget
    sched = make(1, 250, max_batch=8, gb=10, sb=1)
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(5)]
    run(sched)
    assert max(r.batch for r in reqs) == 2 and all(r.state == "done" for r in reqs)


def test_batch_that_cannot_shade_is_retried_one_by_one():
    """A batch whose geometry fits but whose shading cannot, even alone in
    the budget, goes back to the queue and its requests run one by one."""
    sched = make(1, 700, max_batch=4, gb=10, sb=30)   # 2 requests: 200 + 600 > 700
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(2)]
    run(sched)
    assert all(r.state == "done" for r in reqs)
    assert all(r.batch == 1 and r.solo for r in reqs)


def test_kept_workspace_is_reused_and_given_back_under_pressure():
    """Slots keep their memory between jobs: a second job of the same size
    needs nothing new; when a job cannot fit, idle slots give theirs back."""
    sched = make(2, 10**9, keep=True, gb=10, sb=20)
    r0 = sched.submit("a", rays(100), rays(100))
    run(sched)
    assert sched.held_bytes() == 3000 and sched.budget.used == 3000   # geometry + shading
    r1 = sched.submit("a", rays(100), rays(100))
    run(sched)
    assert r0.state == r1.state == "done"
    assert sched.budget.used == 3000 and sched.budget.peak == 3000    # the same slot, no growth
    # slot 1 keeps 3000 bytes: the two jobs cannot both shade in 4500, so one
    # waits (or is preempted) until the other's slot, idle again, gives its
    # memory back
    sched = make(2, 4500, keep=True, gb=10, sb=20)
    st1 = sched.slots[1]
    st1.slot.kept, st1.held = [1000, 2000], 3000
    sched.budget.reserve(3000)
    a = sched.submit("a", rays(100), rays(100))
    b = sched.submit("a", rays(100), rays(100))
    run(sched)
    assert a.state == b.state == "done"
    assert sched.budget.peak <= 4500
    assert sorted(st.held for st in sched.slots) == [0, 3000]
    sched.release_idle()
    assert sched.budget.used == 0


def test_requeued_batch_gives_its_memory_back():
    """A batch that cannot shade is retried one by one; its slot, which kept
    the batch's geometry memory, gives it back, so the single requests are
    not refused for memory the batch left behind."""
    sched = make(2, 1500, max_batch=4, keep=True, gb=3, sb=7)   # per request: 300 + 700
    reqs = [sched.submit("a", rays(100), rays(100)) for _ in range(4)]
    run(sched)
    assert all(r.state == "done" for r in reqs), [(r.state, r.error) for r in reqs]
    assert all(r.solo for r in reqs)
    assert sched.budget.peak <= 1500


class OomSlot(FakeSlot):
    """Fails to allocate the shading of the first `fail` jobs it starts."""

    def __init__(self, fail, **kw):
        super().__init__(**kw)
        self.fail = fail

    def start_shading(self):
        if self.fail > 0:
            self.fail -= 1
            raise SlotOutOfMemory("cached blocks of the wrong shape")
        super().start_shading()


def test_allocation_failure_is_retried_then_rejected():
    """A phase that fails to allocate is retried (the batch one by one, with
    the memory handed back first); a request that keeps failing with
    nothing else in flight is rejected and the rest complete."""
    released = []
    clock = itertools.count()
    sched = Scheduler([OomSlot(fail=2)], HbmBudget(10**9), clock=lambda: next(clock),
                      max_batch=4, on_release=lambda: released.append(1))
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(3)]
    run(sched)
    assert all(r.state == "done" for r in reqs)
    assert all(r.solo for r in reqs) and released
    sched = Scheduler([OomSlot(fail=10**6)], HbmBudget(10**9), clock=lambda: next(clock))
    reqs = [sched.submit("a", rays(10), rays(10)) for _ in range(2)]
    run(sched)
    assert all(r.state == "rejected" and "out of device memory" in r.error for r in reqs)
    assert sched.budget.used == 0
