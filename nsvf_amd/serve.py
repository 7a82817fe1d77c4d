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


class SlotOutOfMemory(RuntimeError):
    """A slot could not allocate its workspace although the budget admitted
    it (the allocator's cached blocks did not fit the request's shape)."""


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
    admitted: Optional[float] = None    # its geometry started (first time)
    shaded: Optional[float] = None      # its shading started
    finished: Optional[float] = None
    result: Any = None
    error: Optional[str] = None
    preemptions: int = 0
    width: int = 0            # image width of the rays (a traversal-order hint), 0 = none
    batch: int = 1            # requests in the job it ran in
    solo: bool = False        # run alone (its batch did not fit)
    oom: int = 0              # allocation failures of its jobs

    @property
    def rays(self) -> int:
        return int(self.ray_start.shape[0])

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


class Job:
    """Requests of one scene rendered as one batch of rays (their rays
    concatenated in order) in one slot.  Every kernel of the path computes
    each ray from its own inputs (the sampler in its evident-intent mode),
    and the hierarchical pass normalises each request's rays over its own
    longest coarse row (the batch's segments), so each request's pixels are
    the same bits as when it runs alone."""

    def __init__(self, requests):
        self.requests = list(requests)
        self.scene = self.requests[0].scene
        self.rid = self.requests[0].rid
        widths = {r.width for r in self.requests}
        # one common image width keeps the traversal's pixel tiles (the
        # concatenation is an image of the same width); with mixed widths the
        # traversal takes each image's pixels in the same tiles through a
        # lane-to-ray order (tiled_order), or row strips when a request has
        # no width -- the results are the same either way
        self.width = widths.pop() if len(widths) == 1 else 0
        self.tiled = self.width == 0 and all(r.width > 0 for r in self.requests)

    def tiled_order(self, device):
        """The concatenated requests' rays in each image's 8x8 pixel tiles
        (an order for Renderer.begin), for a batch of several widths."""
        import torch

        from .render import image_tile_order
        parts, off = [], 0
        for r in self.requests:
            parts.append(image_tile_order(r.rays // r.width, r.width, device) + off)
            off += r.rays
        return torch.cat(parts)

    @property
    def rays(self) -> int:
        return sum(r.rays for r in self.requests)


class Slot:
    """What the scheduler needs from a slot (the GPU one is RenderSlot).
    Byte counts are what the slot would newly allocate (its workspace is
    kept between jobs, so a job that fits what it holds needs 0)."""

    def geometry_bytes(self, job: Job) -> int: ...
    def start_geometry(self, job: Job): ...
    def geometry_ready(self) -> bool: ...
    def shading_bytes(self) -> int: ...
    def start_shading(self): ...
    def shading_done(self) -> bool: ...
    def result(self) -> list: ...          # one result per request of the job
    def reset(self): ...                   # drop the job, keep the workspace
    def workspace_bytes(self) -> int: ...  # bytes the slot holds now
    def release_workspace(self): ...


@dataclass
class _SlotState:
    slot: Slot
    job: Optional[Job] = None
    phase: str = "idle"       # idle, geometry, shading
    held: int = 0             # bytes of the budget the slot holds


class Scheduler:
    """Continuous batching over a fixed set of slots under an HbmBudget.

    poll() advances every slot without blocking: finished shading completes
    its job's requests and frees the slot; finished geometry starts shading
    when its bytes fit; free slots take jobs -- the oldest queued request
    plus up to max_batch - 1 later requests of the same scene among the
    next `window` queued, within max_batch_rays and the budget.  A slot
    keeps its workspace between jobs (no allocation once it has seen the
    largest job); idle slots give theirs back when a job needs the memory."""

    def __init__(self, slots, budget: HbmBudget, clock=time.perf_counter, max_batch=1,
                 max_batch_rays=None, window=16, on_release=None, max_oom_retries=2):
        self.slots = [_SlotState(s) for s in slots]
        self.budget = budget
        self.clock = clock
        self.on_release = on_release    # called after idle slots gave memory back
        self.max_oom_retries = int(max_oom_retries)
        self.max_batch = int(max_batch)
        self.max_batch_rays = max_batch_rays
        self.window = int(window)
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

    def held_bytes(self):
        return sum(s.held for s in self.slots)

    def _true_up(self, st: _SlotState):
        """The budget holds what the slot actually holds."""
        now = int(st.slot.workspace_bytes())
        if now < st.held:
            self.budget.release(st.held - now)
        elif now > st.held:
            self.budget.reserve(now - st.held)
        st.held = now

    def _free_slot(self, st: _SlotState):
        st.slot.reset()
        st.job, st.phase = None, "idle"
        self._true_up(st)

    def _finish(self, st: _SlotState, state, error=None):
        job = st.job
        now = self.clock()
        results = st.slot.result() if state == "done" else [None] * len(job.requests)
        for req, res in zip(job.requests, results):
            req.state, req.error, req.finished = state, error, now
            req.result = res
            req.batch = len(job.requests)
            self.done.append(req)
        self._free_slot(st)

    def _requeue(self, st: _SlotState, solo=False):
        """Puts the slot's job back at the head of the queue (in order); the
        slot gives its workspace back (memory is short when this happens)."""
        for req in reversed(st.job.requests):
            req.state = "queued"
            req.preemptions += 1
            req.solo = req.solo or solo
            self.queue.appendleft(req)
        self._free_slot(st)
        st.slot.release_workspace()
        self._true_up(st)

    def release_idle(self, keep=None):
        """Idle slots (other than `keep`) give their workspace back."""
        freed = False
        for st in self.slots:
            if st is not keep and st.phase == "idle" and st.held > 0:
                st.slot.release_workspace()
                self._true_up(st)
                freed = True
        if freed and self.on_release is not None:
            self.on_release()
        return freed

    def _out_of_memory(self, st: _SlotState, err):
        """A phase failed to allocate: memory goes back to the device and
        the job is retried (its requests one by one if it was a batch);
        a single request that keeps failing with nothing else in flight is
        rejected."""
        retries = max(r.oom for r in st.job.requests) + 1
        for r in st.job.requests:
            r.oom = retries
        others = any(s.phase != "idle" for s in self.slots if s is not st)
        if len(st.job.requests) == 1 and retries > self.max_oom_retries and not others:
            self._finish(st, "rejected", f"out of device memory: {err}")
            self.release_idle()
            return
        self._requeue(st, solo=len(st.job.requests) > 1)
        self.release_idle()
        if self.on_release is not None:
            self.on_release()

    def _take_job(self, st: _SlotState):
        """A job for idle slot st from the queue head, or None if the head
        does not fit now."""
        head = self.queue[0]
        job = Job([head])
        need = st.slot.geometry_bytes(job)
        if not self.budget.fits(need) and self.release_idle(keep=st):
            need = st.slot.geometry_bytes(job)
        if not self.budget.fits(need) and st.held > 0:
            # what this slot keeps may be the wrong shape for the job
            st.slot.release_workspace()
            self._true_up(st)
            need = st.slot.geometry_bytes(job)
        if not self.budget.fits(need):
            return None, need
        taken = [0]
        if not head.solo and self.max_batch > 1:
            for i in range(1, min(len(self.queue), self.window)):
                if len(taken) >= self.max_batch:
                    break
                r = self.queue[i]
                if r.scene != head.scene or r.solo:
                    continue
                if self.max_batch_rays is not None and job.rays + r.rays > self.max_batch_rays:
                    continue
                trial = Job(job.requests + [r])
                n = st.slot.geometry_bytes(trial)
                if not self.budget.fits(n):
                    break
                job, need = trial, n
                taken.append(i)
        for i in reversed(taken):
            del self.queue[i]
        return job, need

    def poll(self) -> bool:
        """One scheduling pass; True if anything changed."""
        changed = False
        # 1. completions free slots
        for st in self.slots:
            if st.phase == "shading" and st.slot.shading_done():
                self._finish(st, "done")
                changed = True
        # 2. geometry done -> shading, if its bytes fit
        for st in self.slots:
            if st.phase == "geometry" and st.slot.geometry_ready():
                need = st.slot.shading_bytes()
                if not self.budget.fits(need):
                    self.release_idle()
                if not self.budget.fits(need) and not any(s.phase == "shading" for s in self.slots):
                    # nothing in flight will free memory: preempt the other
                    # jobs that hold only geometry memory
                    for other in self.slots:
                        if other is not st and other.phase == "geometry":
                            self._requeue(other)
                            changed = True
                    self.release_idle()
                    if not self.budget.fits(need):
                        if len(st.job.requests) > 1:
                            self._requeue(st, solo=True)    # retry its requests one by one
                        else:
                            self._finish(st, "rejected",
                                         f"shading needs {need} bytes; "
                                         f"{self.budget.total - self.budget.used} free with "
                                         f"nothing else in flight")
                        changed = True
                        continue
                if self.budget.fits(need):
                    self.budget.reserve(need)
                    st.held += need
                    now = self.clock()
                    for req in st.job.requests:
                        req.state, req.shaded = "shading", now
                    st.phase = "shading"
                    try:
                        st.slot.start_shading()
                    except SlotOutOfMemory as e:
                        self._out_of_memory(st, e)
                        changed = True
                        continue
                    self._true_up(st)
                    changed = True
        # 3. free slots take jobs (the slot holding the most workspace first)
        idle = sorted((s for s in self.slots if s.phase == "idle"), key=lambda s: -s.held)
        for st in idle:
            if not self.queue:
                break
            job, need = self._take_job(st)
            if job is None:
                if self.in_flight == 0:
                    req = self.queue.popleft()
                    st.job = Job([req])
                    self._finish(st, "rejected",
                                 f"geometry needs {need} bytes; "
                                 f"{self.budget.total - self.budget.used} free")
                    changed = True
                break   # first come, first served: wait for memory
            self.budget.reserve(need)
            st.held += need
            st.job, st.phase = job, "geometry"
            now = self.clock()
            for req in job.requests:
                req.state = "geometry"
                if req.admitted is None:
                    req.admitted = now
            changed = True
            try:
                st.slot.start_geometry(job)
            except SlotOutOfMemory as e:
                self._out_of_memory(st, e)
                break
            self._true_up(st)
        return changed

    def drain(self, sleep=0.0):
        while not self.idle():
            if not self.poll() and sleep:
                time.sleep(sleep)


# ---------------------------------------------------------------------------
# GPU slots and the server


class RenderSlot(Slot):
    """A Renderer on its own stream; its workspace is kept between jobs."""

    def __init__(self, server, index):
        import torch

        from .render import Renderer
        self.server = server
        self.stream = torch.cuda.Stream(device=server.device)   # default priority
        self.renderer = Renderer(server.device, precision=server.precision,
                                 hierarchical=server.hierarchical, fine_steps=server.fine_steps,
                                 compat=server.compat, field_chunk=server.field_chunk)
        self.index = index
        self._done = None
        self._out = None
        self._job = None
        self.tiled_batches = 0   # batches of several image widths traced in pixel tiles
        self._cat_bytes = 0

    def geometry_bytes(self, job):
        sc = self.server.scenes[job.scene]
        R = job.rays
        cat = R * 24 if len(job.requests) > 1 else 0     # the batch's concatenated rays
        return self.renderer.geometry_bytes(sc, R) + cat

    def start_geometry(self, job):
        import torch
        sc = self.server.scenes[job.scene]
        self._job = job
        # the requests' rays were written on the submitting stream
        self.stream.wait_stream(torch.cuda.current_stream(self.server.device))
        try:
            with torch.cuda.stream(self.stream):
                segments = None
                if len(job.requests) == 1:
                    rs, rd = job.requests[0].ray_start, job.requests[0].ray_dir
                else:
                    rs = torch.cat([r.ray_start for r in job.requests])
                    rd = torch.cat([r.ray_dir for r in job.requests])
                    self._cat_bytes = job.rays * 24
                    segments = [r.rays for r in job.requests]
                order = None
                if job.tiled and segments:
                    order = job.tiled_order(rs.device)
                    self.tiled_batches += 1
                self.renderer.begin(sc, rs, rd, job.width, segments=segments, order=order)
        except torch.OutOfMemoryError as e:
            raise SlotOutOfMemory(str(e).splitlines()[0]) from None

    def geometry_ready(self):
        return self.renderer.geometry_ready()

    def shading_bytes(self):
        return self.renderer.shading_bytes()

    def start_shading(self):
        import torch
        try:
            with torch.cuda.stream(self.stream):
                self._out = self.renderer.finish()
                self._done = torch.cuda.Event()
                self._done.record(self.stream)
        except torch.OutOfMemoryError as e:
            raise SlotOutOfMemory(str(e).splitlines()[0]) from None

    def shading_done(self):
        return self._done is not None and self._done.query()

    def result(self):
        rgb, depth, missed = self._out
        out, a = [], 0
        for req in self._job.requests:
            b = a + req.rays
            out.append((rgb[a:b], depth[a:b], missed[a:b]))
            a = b
        return out

    def reset(self):
        pd = self.renderer._pending
        if pd is not None:          # preempted: let its geometry kernels end first
            pd["event"].synchronize()
        self.renderer._pending = None
        self._done = None
        self._out = None
        self._job = None
        self._cat_bytes = 0

    def workspace_bytes(self):
        return self.renderer.workspace_bytes() + self._cat_bytes

    def release_workspace(self):
        self.renderer.release_workspace()


class RenderServer:
    """Resident scenes + a Scheduler over `slots` RenderSlots.

    budget_bytes: device bytes for scenes and request workspaces (default:
    85% of the free device memory at construction).  precision: the field's
    arithmetic (default split BF16, whose per-pixel error against FP64 is
    about 1e-5).  field_chunk: samples per field launch (the split paths'
    workspace grows with it; None = a pass in one launch).  compat: the sampler's
    reference mode (True) or its evident intent (False, the default here): in
    the reference mode a ray whose hit row is full gets one extra sample
    carrying the next ray's voxel id (sample_gpu.cu:188), so a pixel would
    depend on which other pixels share its request."""

    def __init__(self, device, slots=4, budget_bytes=None, precision="bf16x3", hierarchical=True,
                 fine_steps=64.0, compat=False, max_batch=8, max_batch_rays=640_000,
                 field_chunk=1 << 23):
        import torch
        self.device = torch.device(device)
        # samples per field launch: bounds each slot's operand-tile workspace
        # (1 GiB at split BF16 and 32-wide embeddings) whatever the request,
        # and keeps a launch's tiles in the last-level cache between the
        # gather and the MLP
        self.field_chunk = field_chunk
        self.precision = precision
        self.hierarchical = hierarchical
        self.fine_steps = float(fine_steps)
        self.compat = compat
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(free * 0.85)
        self.budget = HbmBudget(budget_bytes)
        self.scenes = {}
        self._scene_bytes = {}
        if compat and max_batch > 1:
            raise ValueError("batching needs the sampler's evident-intent mode (compat=False): "
                             "in the reference's mode a ray reads its neighbours")
        self.scheduler = Scheduler([RenderSlot(self, i) for i in range(slots)], self.budget,
                                   max_batch=max_batch, max_batch_rays=max_batch_rays,
                                   on_release=self._return_cached)

    def _return_cached(self):
        """Workspace the slots gave back goes back to the device (it would
        otherwise stay in the allocator's cache in blocks of the old shapes)."""
        import torch
        torch.cuda.synchronize(self.device)
        torch.cuda.empty_cache()

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
        if any(st.job is not None and st.job.scene == name for st in self.scheduler.slots):
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

    def release_idle(self):
        """Idle slots give their kept workspace back to the allocator."""
        return self.scheduler.release_idle()

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
    requests and of its phases -- queued (arrival to its geometry start),
    geometry (to its shading start: intersection, compaction and the wait
    for shading memory) and shading (to completion) -- and counts."""
    done = [r for r in reqs if r.state == "done"]
    lat = np.array([r.latency for r in done]) * 1e3 if done else np.zeros(1)
    span = (max(r.finished for r in done) - min(r.arrival for r in done)) if done else 0.0
    rays = sum(r.rays for r in done)

    def pct(vals):
        v = np.array(vals) * 1e3 if vals else np.zeros(1)
        return {"p50": round(float(np.percentile(v, 50)), 3),
                "p99": round(float(np.percentile(v, 99)), 3)}

    return {"requests": len(reqs), "done": len(done),
            "rejected": sum(r.state == "rejected" for r in reqs),
            "preemptions": sum(r.preemptions for r in reqs),
            "rays_per_sec": rays / span if span > 0 else None,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "phases_ms": {"queued": pct([r.admitted - r.arrival for r in done]),
                          "geometry": pct([r.shaded - r.admitted for r in done]),
                          "shading": pct([r.finished - r.shaded for r in done])},
            "mean_batch": round(float(np.mean([r.batch for r in done])), 2) if done else 0.0}
