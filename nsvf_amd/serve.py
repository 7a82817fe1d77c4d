"""Serving loop: novel-view requests for several resident scenes rendered
with continuous batching under an HBM budget.

The reference has no server; its closest analogue is the offline free-view
generator (fairnr/renderer.py:125-216: a camera path rendered frame by frame,
one forward per `beam` frames).  Here requests for any resident scene and any
resolution arrive at any time:

  * scenes stay resident: voxel centres, octree, corner embeddings and the
    packed field MLP are uploaded once (DeviceScene) and counted against the
    budget;
  * a request runs in a slot -- a Renderer with its own workspace on its own
    HIP stream (default priority) -- in two phases: geometry (intersection,
    hit compaction, slot offsets; ends with an asynchronous copy of two
    counts to pinned memory) and shading (sampling, field, compositing,
    background).  The scheduler never blocks on the device: it polls each
    slot's events, so phases of different requests overlap on the GPU;
  * admission: a request's geometry starts only when a slot is free and its
    exact workspace bytes fit the budget next to the resident scenes and the
    requests in flight; its shading bytes (known once the geometry is done)
    are admitted the same way.  When a request cannot shade and no shading
    request is in flight to free memory, requests still holding only
    geometry memory are preempted (released and re-queued first); a request
    that cannot fit alone is rejected with RequestTooLarge;
  * each request's result is computed by the same kernels on the same inputs
    as when it runs alone, so it is bitwise the same alone and in the mix.

The scheduler (Scheduler) is plain Python over a small slot interface, so its
admission and ordering logic is exercised on the CPU with stand-in slots.
"""
from __future__ import annotations

import collections
import itertools
import time
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np


class RequestTooLarge(RuntimeError):
    """The request's workspace does not fit the budget even when it runs alone."""


class HbmBudget:
    """Byte accounting of device memory: resident scenes plus the workspace
    of the requests in flight may not exceed `total`."""

    def __init__(self, total_bytes: int):
        self.total = int(total_bytes)
        self.used = 0
        self.peak = 0

    def fits(self, n: int) -> bool:
        return self.used + n <= self.total

    def reserve(self, n: int):
        if not self.fits(n):
            raise RequestTooLarge(f"{n} bytes do not fit ({self.used} of {self.total} in use)")
        self.used += n
        self.peak = max(self.peak, self.used)

    def release(self, n: int):
        self.used -= n
        assert self.used >= 0


@dataclass
class Request:
    rid: int
    scene: str
    ray_start: Any
    ray_dir: Any
    arrival: float
    state: str = "queued"     # queued, geometry, shading, done, rejected
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Any = None
    error: Optional[str] = None
    preemptions: int = 0
    width: int = 0            # image width of the rays (a traversal-order hint), 0 = none

    @property
    def rays(self) -> int:
        return int(self.ray_start.shape[0])

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


class Slot:
    """What the scheduler needs from a slot (the GPU one is RenderSlot)."""

    def geometry_bytes(self, req: Request) -> int: ...
    def start_geometry(self, req: Request): ...
    def geometry_ready(self) -> bool: ...
    def shading_bytes(self) -> int: ...
    def start_shading(self): ...
    def shading_done(self) -> bool: ...
    def result(self) -> Any: ...
    def reset(self): ...


@dataclass
class _SlotState:
    slot: Slot
    req: Optional[Request] = None
    phase: str = "idle"       # idle, geometry, shading
    held: int = 0             # bytes reserved for the request in this slot


class Scheduler:
    """Continuous batching over a fixed set of slots under an HbmBudget.

    poll() advances every slot without blocking: finished shading completes
    its request and frees the slot and its bytes; finished geometry starts
    shading when its bytes fit; free slots take queued requests (first come,
    first served) when their geometry bytes fit."""

    def __init__(self, slots, budget: HbmBudget, clock=time.perf_counter):
        self.slots = [_SlotState(s) for s in slots]
        self.budget = budget
        self.clock = clock
        self.queue = collections.deque()
        self.done = []
        self._ids = itertools.count()

    def submit(self, scene, ray_start, ray_dir, arrival=None, width=0) -> Request:
        req = Request(next(self._ids), scene, ray_start, ray_dir,
                      self.clock() if arrival is None else arrival, width=int(width))
        self.queue.append(req)
        return req

    @property
    def in_flight(self):
        return sum(1 for s in self.slots if s.phase != "idle")

    def idle(self):
        return not self.queue and self.in_flight == 0

    def _finish(self, st: _SlotState, state, error=None):
        req = st.req
        req.state = state
        req.error = error
        req.finished = self.clock()
        if state == "done":
            req.result = st.slot.result()
        self.budget.release(st.held)
        st.slot.reset()
        st.req, st.phase, st.held = None, "idle", 0
        self.done.append(req)

    def _preempt(self, st: _SlotState):
        req = st.req
        req.state = "queued"
        req.preemptions += 1
        self.budget.release(st.held)
        st.slot.reset()
        st.req, st.phase, st.held = None, "idle", 0
        self.queue.appendleft(req)

    def poll(self) -> bool:
        """One scheduling pass; True if anything changed."""
        changed = False
        # 1. completions free slots and bytes
        for st in self.slots:
            if st.phase == "shading" and st.slot.shading_done():
                self._finish(st, "done")
                changed = True
        # 2. geometry done -> shading, if its bytes fit
        for st in self.slots:
            if st.phase == "geometry" and st.slot.geometry_ready():
                need = st.slot.shading_bytes()
                if not self.budget.fits(need) and not any(s.phase == "shading" for s in self.slots):
                    # nothing in flight will free memory: preempt the other
                    # requests that hold only geometry memory
                    for other in self.slots:
                        if other is not st and other.phase == "geometry":
                            self._preempt(other)
                            changed = True
                    if not self.budget.fits(need):
                        self._finish(st, "rejected",
                                     f"shading needs {need} bytes; {self.budget.total - self.budget.used} "
                                     f"free with nothing else in flight")
                        changed = True
                        continue
                if self.budget.fits(need):
                    self.budget.reserve(need)
                    st.held += need
                    st.slot.start_shading()
                    st.phase = "shading"
                    changed = True
        # 3. free slots take queued requests
        for st in self.slots:
            if not self.queue:
                break
            if st.phase != "idle":
                continue
            req = self.queue[0]
            need = st.slot.geometry_bytes(req)
            if not self.budget.fits(need):
                if self.in_flight == 0:
                    self.queue.popleft()
                    st.req = req
                    self._finish(st, "rejected",
                                 f"geometry needs {need} bytes; {self.budget.total - self.budget.used} free")
                    changed = True
                break   # first come, first served: wait for memory
            self.queue.popleft()
            self.budget.reserve(need)
            st.req, st.phase, st.held = req, "geometry", need
            req.state = "geometry"
            if req.admitted is None:
                req.admitted = self.clock()
            st.slot.start_geometry(req)
            changed = True
        return changed

    def drain(self, sleep=0.0):
        while not self.idle():
            if not self.poll() and sleep:
                time.sleep(sleep)


# ---------------------------------------------------------------------------
# GPU slots and the server


class RenderSlot(Slot):
    """A Renderer on its own stream; its workspace lives for one request."""

    def __init__(self, server, index):
        import torch

        from .render import Renderer
        self.server = server
        self.stream = torch.cuda.Stream(device=server.device)   # default priority
        self.renderer = Renderer(server.device, precision=server.precision,
                                 hierarchical=server.hierarchical, fine_steps=server.fine_steps,
                                 compat=server.compat)
        self.index = index
        self._done = None
        self._out = None

    def geometry_bytes(self, req):
        R = req.rays
        return self.renderer.geometry_bytes(self.server.scenes[req.scene], R)

    def start_geometry(self, req):
        import torch
        sc = self.server.scenes[req.scene]
        # the request's rays were written on the submitting stream
        self.stream.wait_stream(torch.cuda.current_stream(self.server.device))
        with torch.cuda.stream(self.stream):
            self.renderer.begin(sc, req.ray_start, req.ray_dir, req.width)

    def geometry_ready(self):
        return self.renderer.geometry_ready()

    def shading_bytes(self):
        return self.renderer.shading_bytes()

    def start_shading(self):
        import torch
        with torch.cuda.stream(self.stream):
            self._out = self.renderer.finish()
            self._done = torch.cuda.Event()
            self._done.record(self.stream)

    def shading_done(self):
        return self._done is not None and self._done.query()

    def result(self):
        return self._out

    def reset(self):
        pd = self.renderer._pending
        if pd is not None:          # preempted: let its geometry kernels end first
            pd["event"].synchronize()
        self.renderer.release_workspace()
        self.renderer._pending = None
        self._done = None
        self._out = None


class RenderServer:
    """Resident scenes + a Scheduler over `slots` RenderSlots.

    budget_bytes: device bytes for scenes and request workspaces (default:
    90% of the free device memory at construction).  precision: the field's
    arithmetic (default split BF16, whose per-pixel error against FP64 is
    about 1e-5).  compat: the sampler's
    reference mode (True) or its evident intent (False, the default here): in
    the reference mode a ray whose hit row is full gets one extra sample
    carrying the next ray's voxel id (sample_gpu.cu:188), so a pixel would
    depend on which other pixels share its request."""

    def __init__(self, device, slots=4, budget_bytes=None, precision="bf16x3", hierarchical=True,
                 fine_steps=64.0, compat=False):
        import torch
        self.device = torch.device(device)
        self.precision = precision
        self.hierarchical = hierarchical
        self.fine_steps = float(fine_steps)
        self.compat = compat
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(free * 0.9)
        self.budget = HbmBudget(budget_bytes)
        self.scenes = {}
        self._scene_bytes = {}
        self.scheduler = Scheduler([RenderSlot(self, i) for i in range(slots)], self.budget)

    def load_scene(self, name, scene):
        """Uploads a VoxelScene (or takes a DeviceScene) and keeps it resident."""
        from .render import DeviceScene
        if name in self.scenes:
            self.unload_scene(name)
        dsc = scene if isinstance(scene, DeviceScene) else DeviceScene(scene, self.device)
        n = dsc.nbytes()
        self.budget.reserve(n)
        self.scenes[name] = dsc
        self._scene_bytes[name] = n
        return dsc

    def unload_scene(self, name):
        if any(st.req is not None and st.req.scene == name for st in self.scheduler.slots):
            raise RuntimeError(f"scene {name} has requests in flight")
        del self.scenes[name]
        self.budget.release(self._scene_bytes.pop(name))

    def submit(self, scene, ray_start, ray_dir, arrival=None, width=0):
        """Queues a view: rays [R,3] on the device; width as in
        Renderer.begin()."""
        if scene not in self.scenes:
            raise KeyError(f"scene {scene} is not resident")
        return self.scheduler.submit(scene, ray_start, ray_dir, arrival, width)

    def poll(self):
        return self.scheduler.poll()

    def drain(self):
        self.scheduler.drain()

    def run_arrivals(self, arrivals):
        """Replays (offset_seconds, scene, ray_start, ray_dir[, width])
        arrivals against the wall clock -- each submitted when its time comes,
        while earlier requests are in flight -- and returns the requests when
        all are done."""
        arrivals = sorted(arrivals, key=lambda a: a[0])
        t0 = time.perf_counter()
        reqs = []
        i = 0
        while i < len(arrivals) or not self.scheduler.idle():
            now = time.perf_counter() - t0
            while i < len(arrivals) and arrivals[i][0] <= now:
                off, scene, rs, rd, *w = arrivals[i]
                reqs.append(self.submit(scene, rs, rd, arrival=t0 + off, width=w[0] if w else 0))
                i += 1
            self.poll()
        return reqs


def latency_summary(reqs):
    """rays/s over the span of the run, p50/p99 latency (ms) of the done
    requests, and counts."""
    done = [r for r in reqs if r.state == "done"]
    lat = np.array([r.latency for r in done]) * 1e3 if done else np.zeros(1)
    span = (max(r.finished for r in done) - min(r.arrival for r in done)) if done else 0.0
    rays = sum(r.rays for r in done)
    return {"requests": len(reqs), "done": len(done),
            "rejected": sum(r.state == "rejected" for r in reqs),
            "preemptions": sum(r.preemptions for r in reqs),
            "rays_per_sec": rays / span if span > 0 else None,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))}
