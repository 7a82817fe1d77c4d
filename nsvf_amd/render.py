"""Fast view renderer over the fused HIP pipeline, with the scene resident on
the device.

One call renders a batch of rays (a view, or several requests' rays) through
the stages of fairnr/models/fairnr_model.py:142-186 for the NSVF model with
hierarchical sampling:

  intersect (SVO) -> sort -> hit compaction -> coarse inverse-CDF samples ->
  field (gather + MLP) -> composite -> fine inverse-CDF samples over the
  coarse samples -> field -> composite -> background fill.

All stages are kernels of libnsvf_amd.so on the current stream.  One device
to host read per view (hit-ray count and coarse slot total) sizes the sample
streams; the reference reads back far more (hit compaction, max lengths, a
per-column sum in its renderer loop).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import geometry
from ._lib import check, lib
from .field import pack_mlp
from .scene import VoxelScene

BG_DEPTH = 5.0              # fairnr/modules/implicit.py:19
BG_COLOR = (1.0, 1.0, 1.0)  # fairnr/modules/field.py:69 (transparent background)
# field MLP arithmetic: BF16 MFMA, FP32 MFMA, or split BF16 (three BF16 MFMAs
# per product on hi / lo operand parts, about 16 significant bits)
PRECISIONS = {"bf16": 0, "fp32": 1, "bf16x3": 2}
_WEIGHTS = {0: "bf16", 1: "f32", 2: "bf16x3"}


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def node_boxes(centers, children, half_voxel):
    """Per-node boxes of the octree: leaf boxes are fp32(c - h), fp32(c + h);
    an internal node's box is the bounding box of its children's boxes
    (children have smaller ids than their parent, so one pass per depth from
    the leaves up fills it).  Returns mn [T,3], mx [T,3] f32, scale [T] i32."""
    h = np.float32(half_voxel)
    scale = children[:, 8].astype(np.int32)
    T = children.shape[0]
    mn = np.zeros((T, 3), np.float32)
    mx = np.zeros((T, 3), np.float32)
    leaf = scale == 1
    mn[leaf] = (centers[leaf] - h).astype(np.float32)
    mx[leaf] = (centers[leaf] + h).astype(np.float32)
    kids = children[:, :8].astype(np.int32)
    for sc in np.unique(scale[~leaf]):          # increasing depth order from the leaves up
        nodes = np.nonzero(scale == sc)[0]
        k = kids[nodes]                           # [n,8]
        valid = k >= 0
        kk = np.where(valid, k, 0)
        cmn = np.where(valid[..., None], mn[kk], np.float32(np.inf))
        cmx = np.where(valid[..., None], mx[kk], np.float32(-np.inf))
        mn[nodes] = cmn.min(1)
        mx[nodes] = cmx.max(1)
    return mn, mx, scale


def pass_seed(seed, k):
    """Noise seed of sampling pass k (0 coarse, 1 fine) of a view rendered
    with `seed` (None: -1, no noise): seed + k * 0x9E3779B9 mod 2^32, so the
    two passes draw unrelated streams."""
    if seed is None:
        return -1
    seed = int(seed)
    if not 0 <= seed < 1 << 32:
        raise ValueError(f"noise seed {seed}: not in [0, 2^32)")
    return (seed + k * 0x9E3779B9) & 0xFFFFFFFF


def tree_depth(children):
    """Depth of the octree whose root is the last node: its children[8] is
    1 << (depth + 1) (octree.cpp:119); a tree of one leaf has depth -1."""
    scale = int(children[-1, 8])
    if scale < 1 or scale & (scale - 1):
        raise ValueError(f"octree root scale {scale} is not a power of two")
    return scale.bit_length() - 2


def pack_tree(centers, children, half_voxel):
    """The octree packed as node records for the traversal: record k holds,
    for child slot s, (min xyz, scale bits) and (max xyz, id bits) of child
    children[k][s] (id -1 for an empty slot); record T holds the root (node
    T-1) in slot 0.  Returns float32 [T + 1, 16, 4].

    A child with exactly one child of its own stands for that child, and so
    on down the chain: a one-child node's exact box is its child's box, so
    the walk tests the same boxes and reaches the same leaves in the same
    order, one dependent record fetch per chain link fewer.  The reference's
    octree has a whole level of them: voxel centres sit on even half-voxel
    coordinates, so the finest internal nodes each hold one voxel.  Records
    of nodes that only chains reach stay in place, unread."""
    T = children.shape[0]
    mn, mx, scale = node_boxes(centers, children, half_voxel)
    kids = children[:, :8].astype(np.int32)
    only = np.where((kids >= 0).sum(1) == 1, kids.max(1), -1)
    ids = np.full((T + 1, 8), -1, np.int32)
    ids[:T] = kids
    ids[T, 0] = T - 1
    # children have smaller ids than their parents, so every chain ends
    while True:
        nxt = np.where(ids >= 0, only[np.maximum(ids, 0)], -1)
        if not (nxt >= 0).any():
            break
        ids = np.where(nxt >= 0, nxt, ids)
    ok = ids >= 0
    k = np.where(ok, ids, 0)
    rec = np.zeros((T + 1, 8, 2, 4), np.float32)
    rec[:, :, 0, :3] = np.where(ok[..., None], mn[k], np.float32(0))
    rec[:, :, 0, 3] = np.where(ok, scale[k], 0).astype(np.int32).view(np.float32)
    rec[:, :, 1, :3] = np.where(ok[..., None], mx[k], np.float32(0))
    rec[:, :, 1, 3] = ids.view(np.float32)
    return rec.reshape(T + 1, 16, 4)


def coherent_order(ray_start, ray_dir):
    """A traversal grouping for rays that are not image rows: int32 [R] on
    the rays' device, the ray indices sorted by direction octant (the union
    walk's child order), then origin cell (a 16^3 grid over the origins'
    bounds), then the Morton code of the direction (10 bits per axis), so
    that each wave's 64 rays are close in origin and direction
    (nsvf_ray_order_keys, one pass, then a sort).  Any order gives the same
    hits (nsvf_render_hits_ordered)."""
    rs = ray_start.contiguous().float()
    rd = ray_dir.contiguous().float()
    R = rs.shape[0]
    keys = torch.empty(R, dtype=torch.int64, device=rs.device)
    if R == 0:
        return keys.to(torch.int32)
    lo, hi = torch.aminmax(rs, dim=0)
    bounds = torch.cat([lo, hi])
    check(lib().nsvf_ray_order_keys(_p(rs), _p(rd), R, _p(bounds), _p(keys), _stream(rs.device)),
          "ray_order_keys")
    return torch.argsort(keys).to(torch.int32)


_TILE_ORDERS = {}


def image_tile_order(h, w, device):
    """The ray indices of an h x w image (row-major rays) in 8x8 pixel
    tiles -- tiles row by row, pixels row by row within a tile -- as int32 on
    `device` (cached per size): the grouping the traversal uses for a width
    hint, as an order for images batched with others of other widths."""
    key = (int(h), int(w), str(device))
    t = _TILE_ORDERS.get(key)
    if t is None:
        i = np.arange(h * w, dtype=np.int64)
        y, x = i // w, i % w
        k = ((y // 8) * ((w + 7) // 8) + x // 8) * 64 + (y % 8) * 8 + (x % 8)
        t = torch.from_numpy(np.argsort(k, kind="stable").astype(np.int32)).to(device)
        _TILE_ORDERS[key] = t
    return t


class DeviceScene:
    """A voxel scene resident on one GPU: shifted voxel centres (gather),
    octree over the unshifted centres (intersection), corner ids, embeddings
    and the packed field MLP.  Built once; rendering reads it only."""

    def __init__(self, scene: VoxelScene, device):
        self.scene = scene
        self.device = torch.device(device)
        vs = np.float32(scene.voxel_size)
        pts = torch.from_numpy(scene.points)
        # encoder.py:383: the gather (and brute-force intersection) use centres
        # shifted by voxel_size / 10 in x; the octree is built unshifted.
        shifted = pts.clone()
        shifted[:, 0] += torch.tensor(vs) / 10
        cen, ch = geometry.build_easy_octree(pts, torch.tensor(vs) / 2)
        self.centers = shifted.contiguous().to(self.device)
        # the octree packed for the traversal (boxes and child ids)
        rec = pack_tree(cen.float().numpy(), ch.int().numpy(), np.float32(vs) * np.float32(0.5))
        self.tree_records = torch.from_numpy(rec).to(self.device)
        self.tree_depth = tree_depth(ch.int().numpy())
        self.feats = torch.from_numpy(scene.feats.astype(np.int32)).to(self.device)
        self.values = torch.from_numpy(scene.values).to(self.device)
        self.voxel_size = float(vs)
        self.step_size = float(np.float32(scene.step_size))
        self.max_hits = int(scene.max_hits)
        self.mlp = pack_mlp(scene.mlp, self.device)
        self.D, self.H, self.L = self.mlp["shape"]
        # the gather backward's corner-incidence order (training), made on
        # first use
        self.gather_plan = None
        if not lib().nsvf_field_supported(self.D, self.H, self.L):
            raise RuntimeError(f"field shape D={self.D} H={self.H} L={self.L} not supported")

    @property
    def num_nodes(self):
        return self.tree_records.shape[0] - 1

    def nbytes(self):
        ts = [self.centers, self.tree_records,
              self.feats, self.values,
              self.mlp["bf16"], self.mlp["bf16x3"], self.mlp["f32"], self.mlp["bias"]]
        return sum(t.numel() * t.element_size() for t in ts)


@dataclass
class CoarseStream:
    """Output of the geometry stages for R rays: sorted hit rows, the list of
    the n_hit rays that hit a voxel, and the coarse sample stream -- `slots`
    16-byte records {voxel id (-1 = unused), depth, interval, ray id}; ray r
    owns slots [off[r], off[r] + bound) of which the first cnt[r] are
    samples.  `scal` holds the device copies of [n_hit, slots, longest row,
    fine slots]."""
    rays: int
    n_hit: int
    slots: int
    hits: torch.Tensor
    count: torch.Tensor
    tot: torch.Tensor
    flag: torch.Tensor
    rank: torch.Tensor
    hit_list: torch.Tensor
    off: torch.Tensor
    cnt: torch.Tensor
    samples: torch.Tensor
    f_bound: torch.Tensor
    scal: torch.Tensor

    def owned(self):
        """A copy whose tensors no later coarse_stream call overwrites (what
        an autograd graph keeps until its backward)."""
        c = lambda t: t.clone()
        return CoarseStream(self.rays, self.n_hit, self.slots, None, None, None, c(self.flag), None,
                            c(self.hit_list), c(self.off), c(self.cnt), c(self.samples), None,
                            c(self.scal))

    @property
    def c_off(self):
        return self.off

    @property
    def c_cnt(self):
        return self.cnt


@dataclass
class RenderStats:
    rays: int
    hit_rays: int
    coarse_slots: int     # slots of the coarse stream (samples + unused)
    fine_slots: int       # upper bound of the fine stream's slots
    coarse_width: int


class Renderer:
    """Renders rays against a DeviceScene.  Workspace buffers are kept and
    grown as needed, so repeated views of the same size allocate nothing."""

    def __init__(self, device, precision="bf16", hierarchical=True, fine_steps=64.0,
                 compat=True, split_field=None, reuse_coarse=True, fold_composite=True,
                 noise_seed=None, tolerance=0.0, chunk_size=65536, field_chunk=None):
        self.device = torch.device(device)
        # the split field paths stage operand tiles for every sample of a
        # launch (128 B per sample for split BF16); field_chunk (samples)
        # bounds that workspace by evaluating a pass in launches of at most
        # that many samples -- same values, fewer bytes held (serving)
        if field_chunk is not None and int(field_chunk) < 1024:
            raise ValueError(f"field_chunk {field_chunk}: at least 1024 samples")
        self.field_chunk = None if field_chunk is None else int(field_chunk)
        # early ray termination (--raymarching-tolerance, renderer.py:136-174):
        # > 0 keeps, in each pass, the field of exactly the samples the
        # reference's marching in chunks of at most chunk_size rows and hits
        # evaluates (rays stop once their free energy passes -log(tolerance))
        # and zeroes the others; 0 (the reference's default) keeps them all
        if not tolerance >= 0.0 or tolerance >= 1.0:
            raise ValueError(f"tolerance {tolerance}: in [0, 1)")
        if int(chunk_size) <= 0 or int(chunk_size) >= 1 << 31:
            raise ValueError(f"chunk_size {chunk_size}: in [1, 2^31)")
        self.tolerance = float(tolerance)
        self.chunk_size = int(chunk_size)
        # None: every step at the middle of its CDF interval (noise 0.5, the
        # evaluation sampler); an int in [0, 2^32): training-mode noise of the
        # samplers (clib/__init__.py:250-256), counter-based, a different
        # stream per seed and pass
        self.noise_seed = noise_seed
        # hierarchical views: the fine sampler composites the coarse rows
        # itself (the coarse composite's weights, bit for bit) instead of a
        # coarse composite kernel whose colour the fine pass replaces
        self.fold_composite = bool(fold_composite)
        # fine pass: copy the coarse field value of every fine sample that is
        # its bin's coarse sample again (same voxel and depth bits) and
        # evaluate the field on the other fine samples only (same results)
        self.reuse_coarse = bool(reuse_coarse)
        # BF16: gather kernel + MLP kernel through a workspace (True), one
        # fused kernel (False), or per scene (None): fused where the fused
        # kernel stages the field's weights in LDS, else split
        self.split_field = split_field
        if precision not in PRECISIONS:
            raise ValueError(f"precision {precision!r}: one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.hierarchical = hierarchical
        self.fine_steps = float(fine_steps)
        self.compat = int(bool(compat))
        self._ws = {}
        self._host = None      # pinned [n_hit, slots] of the view in flight
        self._seg_host = None  # pinned segment bounds of a batch of views
        self._pending = None   # the view between begin() and finish()
        self.last = None
        self.trace = None   # optional list of (stage, event) marks
        self._kernel_events = None  # optional list of (kernel, units, start, end)
        # optional float32 [num_voxels] running max of per-ray voxel
        # probabilities, updated by every pass (encoder.py:594-603)
        self.track = None

    def record_kernels(self, on=True):
        """Start (or stop) recording events around the field-kernel launches;
        returns the list the records go to."""
        self._kernel_events = [] if on else None
        return self._kernel_events

    def _buf(self, name, n, dtype):
        t = self._ws.get(name)
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(self._alloc_elems(n), dtype=dtype, device=self.device)
            self._ws[name] = t
        return t[:n] if n else t[:0]

    def _mark(self, stage):
        if self.trace is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self.trace.append((stage, e))

    def _stream(self, name, m):
        """m sample records {idx, depth, dist, ray} as an int32 [m, 4] buffer."""
        return self._buf(name, m * 4, torch.int32).view(m, 4)

    def _track(self, sc, hit_list, scal, n_hit, off, cnt, samples, probs, st):
        """Folds a pass's sample probabilities into self.track (renderer.py:
        250-251 calls track_voxel_probs after each pass when tracking)."""
        if self.track is None or n_hit <= 0:
            return
        t = self.track
        if t.dtype != torch.float32 or t.device != self.device or t.numel() != sc.scene.num_voxels:
            raise ValueError("track must be a float32 [num_voxels] tensor on the renderer's device")
        check(lib().nsvf_track_voxel_probs(_p(hit_list), _p(scal[0:1]), n_hit, _p(off), _p(cnt),
                                           _p(samples), _p(probs), _p(t), t.numel(), st),
              "track_voxel_probs")

    def uses_split_field(self, sc):
        """Whether the BF16 field of scene `sc` runs as gather + MLP kernels."""
        if self.split_field is not None:
            return bool(self.split_field)
        return not lib().nsvf_field_fused_supported(sc.D, sc.H, sc.L)

    def _field(self, sc, ray_start, ray_dir, samples, m_dev, m, out, slots=None):
        """Field of the m_dev samples of a stream (slots: optional evaluation
        list of stream slots; then only those slots of out are written)."""
        L = lib()
        prec = PRECISIONS[self.precision]
        w = sc.mlp[_WEIGHTS[prec]]
        st = _stream(self.device)
        ev = self._kernel_events is not None and m > 0

        def event():
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            return e

        if prec == 2 or (prec == 0 and self.uses_split_field(sc)):
            embed, mlp = ((L.nsvf_field_embed_bf16x3, L.nsvf_field_mlp_bf16x3) if prec == 2 else
                          (L.nsvf_field_embed_bf16, L.nsvf_field_mlp_bf16))
            C = m if self.field_chunk is None else min(m, self.field_chunk)
            ws_bytes = int(L.nsvf_field_workspace_bytes(prec, sc.D, sc.H, sc.L, C))
            ws = self._buf("field_ws", (ws_bytes + 15) // 16 * 4, torch.int32)
            if C < m:
                # launch k takes samples [kC, kC + C) of the m_dev: its count
                # clamp(m_dev - kC, 0, C) computed on the device
                K = -(-m // C)
                ar = torch.arange(K, dtype=torch.int64, device=self.device) * C
                cnts = (m_dev[:1].to(torch.int64) - ar).clamp_(0, C).to(torch.int32)
            else:
                K, cnts = 1, None
            for k in range(K):
                mk = m_dev if cnts is None else cnts[k:k + 1]
                if slots is not None:
                    lst, smp, o = slots[k * C:], samples, out
                else:
                    lst, smp, o = None, samples[k * C:], out[k * C * 4:]
                e0 = event() if ev else None
                check(embed(sc.D, sc.H, sc.L, _p(sc.centers), _p(sc.feats), _p(sc.values),
                            sc.voxel_size, _p(ray_start), _p(ray_dir), _p(smp), _p(lst),
                            _p(mk), C, _p(ws), ws_bytes, st), "field_embed")
                e1 = event() if ev else None
                check(mlp(sc.D, sc.H, sc.L, _p(ws), ws_bytes, _p(mk), C, _p(w),
                          _p(sc.mlp["bias"]), _p(o), _p(lst), st), "field_mlp")
                if ev:
                    e2 = event()
                    self._kernel_events.append(("embed", C, e0, e1))
                    self._kernel_events.append(("mlp", C, e1, e2))
            return
        e0 = event() if ev else None
        check(L.nsvf_field_eval(prec, sc.D, sc.H, sc.L, _p(sc.centers), _p(sc.feats), _p(sc.values),
                                sc.voxel_size, _p(ray_start), _p(ray_dir), _p(samples), _p(slots),
                                _p(m_dev), m, _p(w), _p(sc.mlp["bias"]), _p(out), None, 0, st),
              "field_eval")
        if ev:
            self._kernel_events.append(("field", m, e0, event()))

    def _reuse(self):
        """Fine-pass field reuse (off under early termination: the marched
        chunks decide which fine samples are evaluated)."""
        return self.reuse_coarse and self.tolerance <= 0.0

    def _march(self, samples, slots, hit_list, n_hit, off, cnt, S, out, tag):
        """Early ray termination over a pass whose field `out` holds every
        sample's value: replays the reference's chunked marching (chunks of
        at most chunk_size rows and hits, csrc/march.hip), marks the samples
        it evaluates in the `{tag}_eval` buffer and zeroes the field of the
        others; `{tag}_active` ends nonzero for the rows still marching.  off /
        cnt: the pass's frame-indexed slot offsets and counts; S: its
        longest row."""
        L = lib()
        i32 = torch.int32
        st = _stream(self.device)
        mark = self._buf(f"{tag}_eval", max(slots, 1), torch.int8)
        mark.zero_()
        active = self._buf(f"{tag}_active", max(n_hit, 1), i32)
        active.fill_(1)
        if n_hit <= 0 or S <= 0:
            return
        chunk = self.chunk_size
        G = -(-n_hit // chunk)
        hist = self._buf("m_hist", int(L.nsvf_march_hist_ints(n_hit, chunk, S)), i32)
        hist.zero_()
        span = self._buf("m_span", 2 * G, i32)
        nxt = self._buf("m_next", G, i32)
        nxt[:G].zero_()
        total = self._buf("m_total_fe", n_hit, torch.float32)
        thr = float(np.float32(-math.log(self.tolerance)))
        check(L.nsvf_march_begin(_p(hit_list), n_hit, _p(cnt), chunk, S, _p(active), _p(total),
                                 _p(hist), st), "march_begin")
        # at most S chunks per group; the host looks every 16 steps whether
        # every group is done
        for step in range(S):
            check(L.nsvf_march_step(_p(hit_list), n_hit, _p(off), _p(cnt), _p(samples), _p(out),
                                    chunk, S, thr, _p(active), _p(total), _p(hist), _p(span),
                                    _p(nxt), _p(mark), st), "march_step")
            if step % 16 == 15 and int(nxt[:G].min()) >= S:
                break
        check(L.nsvf_march_end(_p(hit_list), n_hit, _p(off), _p(cnt), _p(mark), _p(out), st),
              "march_end")

    def _scan(self, counts, offsets, n, total):
        ws = self._buf("scan_ws", int(lib().nsvf_scan_workspace_ints(n)), torch.int32)
        check(lib().nsvf_scan_exclusive_i32(_p(counts), _p(offsets), n, _p(ws), _p(total),
                                            _stream(self.device)), "scan")

    # ---- workspace planning (exact bytes each phase allocates)

    @staticmethod
    def _alloc_elems(n):
        return max(int(n * 1.25), 1024)

    def _plan_bytes(self, sizes):
        """Bytes that _buf would newly allocate for {name: (n, dtype)}."""
        total = 0
        for name, (n, dtype) in sizes.items():
            t = self._ws.get(name)
            if t is None or t.numel() < n or t.dtype != dtype:
                total += self._alloc_elems(n) * torch.empty((), dtype=dtype).element_size()
        return total

    def workspace_bytes(self):
        return sum(t.numel() * t.element_size() for t in self._ws.values())

    def release_workspace(self):
        """Drops every workspace buffer (back to torch's caching allocator)."""
        self._ws = {}

    def _geometry_sizes(self, R, P):
        i32, f32 = torch.int32, torch.float32
        return {"hits": (R * P * 4, i32), "count": (R, i32), "tot": (R, f32), "flag": (R, i32),
                "c_bound": (R, i32), "scalars": (8, i32), "f_bound": (R, i32), "rank": (R, i32),
                "hit_list": (R, i32), "c_off": (R, i32),
                "scan_ws": (int(lib().nsvf_scan_workspace_ints(R)), i32),
                "ray_rgb": (R * 3, f32), "ray_depth": (R, f32), "ray_missed": (R, f32),
                "c_cnt": (R, i32), "f_off": (R, i32), "f_cnt": (R, i32),
                "out": (R * 5, f32), "seg_bounds": (17, i32), "seg_max": (16, i32)}

    def _shading_sizes(self, sc, R, n_hit, slots_c, slots_f):
        i32, f32 = torch.int32, torch.float32
        sizes = {"c_samples": (slots_c * 4, i32), "c_field": (slots_c * 4, f32),
                 "c_prob": (slots_c, f32)}
        m = slots_c
        if self.hierarchical and n_hit > 0:
            sizes.update({"f_samples": (slots_f * 4, i32), "f_field": (slots_f * 4, f32)})
            if self._reuse():
                nb = int(lib().nsvf_render_eval_list_blocks(slots_f))
                sizes.update({"f_src": (slots_f, torch.int16), "e_bcnt": (nb, i32),
                              "e_boff": (nb, i32), "e_list": (slots_f, i32),
                              "scan_ws": (int(lib().nsvf_scan_workspace_ints(max(R, nb))), i32)})
            m = max(m, slots_f)
        prec = PRECISIONS[self.precision]
        if (prec == 2 or (prec == 0 and self.uses_split_field(sc))) and m > 0:
            if self.field_chunk is not None:
                m = min(m, self.field_chunk)
            ws_bytes = int(lib().nsvf_field_workspace_bytes(prec, sc.D, sc.H, sc.L, m))
            sizes["field_ws"] = ((ws_bytes + 15) // 16 * 4, i32)
        return sizes

    def geometry_bytes(self, sc: DeviceScene, R):
        """Workspace bytes begin() would allocate for R rays."""
        return self._plan_bytes(self._geometry_sizes(R, sc.max_hits))

    def shading_bytes(self):
        """Workspace bytes finish() would allocate for the view begun last
        (valid once geometry_ready())."""
        pd = self._pending
        n_hit, slots_c, slots_f = self._counts()
        return self._plan_bytes(self._shading_sizes(pd["sc"], pd["R"], n_hit, slots_c, slots_f))

    # ---- phase 1: intersection, compaction, slot offsets

    @torch.no_grad()
    def begin(self, sc: DeviceScene, ray_start, ray_dir, width=0, segments=None, order=None):
        """Enqueues the geometry stages of a view on the current stream:
        intersection, hit compaction and the coarse slot offsets
        (encoder.py:498-535), then an asynchronous copy of the hit-ray count
        and slot total to pinned host memory.  geometry_ready() tells when
        finish() can run without waiting.  width: 0, or the pixel width of
        the image(s) whose rows the rays are (a traversal-order hint; the
        results do not depend on it).  segments: for several views rendered
        as one batch, their ray counts in order (at most 16; the rays are
        their concatenation); each view's hierarchical pass then normalises
        over its own longest coarse row, so every view gets the bits it
        gets alone (with the sampler's evident-intent mode).  order: instead
        of width, the traversal's grouping of the rays (int32 [R] device
        tensor, a permutation -- e.g. coherent_order(ray_start, ray_dir) --
        for rays that are not image rows; the results do not depend on
        it)."""
        ray_start = ray_start.contiguous()
        ray_dir = ray_dir.contiguous()
        R = ray_start.shape[0]
        if width and R % width:
            raise ValueError(f"{R} rays are not rows of width {width}")
        if segments is not None:
            segments = [int(n) for n in segments]
            if sum(segments) != R or len(segments) > 16 or min(segments) < 0:
                raise ValueError(f"segments {segments} do not split {R} rays into <= 16 views")
            if len(segments) == 1:
                segments = None

        if order is not None:
            order = order.contiguous()
            if order.dtype != torch.int32 or order.shape != (R,) or order.device != ray_start.device:
                raise ValueError(f"order must be an int32 [{R}] tensor on the rays' device")

        def intersect(sc, P, hits, count, tot, flag, c_bound, st):
            if order is not None:
                check(lib().nsvf_render_hits_ordered(_p(ray_start), _p(ray_dir), R, _p(order),
                                                     _p(sc.tree_records), sc.num_nodes,
                                                     sc.tree_depth, P, sc.step_size, _p(hits),
                                                     _p(count), _p(tot), _p(flag), _p(c_bound),
                                                     st), "render_hits_ordered")
                return
            check(lib().nsvf_render_hits(_p(ray_start), _p(ray_dir), R, int(width),
                                         _p(sc.tree_records), sc.num_nodes, sc.tree_depth, P,
                                         sc.step_size,
                                         _p(hits), _p(count), _p(tot), _p(flag), _p(c_bound), st),
                  "render_hits")
        self._begin(sc, R, ray_start, ray_dir, intersect, segments)

    @torch.no_grad()
    def begin_camera(self, sc: DeviceScene, camera):
        """begin() for the rays of a camera (nsvf_amd.rays.ViewCamera): they
        are generated inside the traversal kernel (renderer.py:91-112,
        reader.py:92-119) into the workspace, in row-major pixel order."""
        R = camera.num_rays
        rs = self._buf("cam_rs", R * 3, torch.float32).view(R, 3)
        rd = self._buf("cam_rd", R * 3, torch.float32).view(R, 3)
        cam = camera.c()

        def intersect(sc, P, hits, count, tot, flag, c_bound, st):
            check(lib().nsvf_render_hits_camera(ctypes.byref(cam), _p(sc.tree_records),
                                                sc.num_nodes, sc.tree_depth, P, sc.step_size,
                                                _p(rs), _p(rd),
                                                _p(hits), _p(count), _p(tot), _p(flag),
                                                _p(c_bound), st), "render_hits_camera")
        self._begin(sc, R, rs, rd, intersect)

    def _begin(self, sc, R, ray_start, ray_dir, intersect, segments=None):
        L = lib()
        dev = self.device
        st = _stream(dev)
        P = sc.max_hits
        for name, (n, dtype) in self._geometry_sizes(R, P).items():
            self._buf(name, n, dtype)
        if segments is not None:
            # segment bounds to the device through pinned memory (the previous
            # view's copy has completed: a renderer runs one view at a time)
            if self._seg_host is None:
                self._seg_host = torch.zeros(17, dtype=torch.int32, pin_memory=True)
            self._seg_host[: len(segments) + 1] = torch.from_numpy(
                np.concatenate([[0], np.cumsum(segments)]).astype(np.int32))
            self._buf("seg_bounds", 17, torch.int32).copy_(self._seg_host, non_blocking=True)
            self._buf("seg_max", 16, torch.int32).zero_()
        self._mark("start")
        hits = self._buf("hits", R * P * 4, torch.int32)
        count = self._buf("count", R, torch.int32)
        tot = self._buf("tot", R, torch.float32)
        flag = self._buf("flag", R, torch.int32)
        c_bound = self._buf("c_bound", R, torch.int32)
        intersect(sc, P, hits, count, tot, flag, c_bound, st)
        self._mark("intersect")
        scal = self._buf("scalars", 8, torch.int32)   # [n_hit, slots_c, max_c, slots_f]
        scal.zero_()
        f_bound = self._buf("f_bound", R, torch.int32)
        f_bound.zero_()
        rank = self._buf("rank", R, torch.int32)
        self._scan(flag, rank, R, scal[0:1])
        hit_list = self._buf("hit_list", R, torch.int32)
        check(L.nsvf_render_compact(_p(flag), _p(rank), R, _p(hit_list), st), "render_compact")
        c_off = self._buf("c_off", R, torch.int32)
        self._scan(c_bound, c_off, R, scal[1:2])
        if self._host is None:
            self._host = torch.zeros(2, dtype=torch.int32, pin_memory=True)
        # the one device-to-host read of the view: hit rays and slot total
        self._host.copy_(scal[:2], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._pending = {"sc": sc, "rs": ray_start, "rd": ray_dir, "R": R, "event": ev,
                         "segments": segments}
        self._mark("compact")

    def geometry_ready(self):
        return self._pending is not None and self._pending["event"].query()

    def _counts(self):
        self._pending["event"].synchronize()
        n_hit, slots_c = int(self._host[0]), int(self._host[1])
        slots_f = slots_c + n_hit * (int(math.ceil(self.fine_steps)) + 2)   # fine-slot bound
        return n_hit, slots_c, slots_f

    def _segment_args(self):
        segs = self._pending.get("segments")
        if not segs:
            return (None, 0, None)
        return (_p(self._buf("seg_bounds", 17, torch.int32)), len(segs),
                _p(self._buf("seg_max", 16, torch.int32)))

    def _pass_seed(self, k):
        """The sampler seed of pass k (0 coarse, 1 fine): -1 without noise."""
        return pass_seed(self.noise_seed, k)

    def _sample_coarse(self):
        """Coarse sample stream of the view begun last (its geometry done)."""
        L = lib()
        pd = self._pending
        sc, R = pd["sc"], pd["R"]
        n_hit, slots_c, slots_f = self._counts()
        for name, (n, dtype) in self._shading_sizes(sc, R, n_hit, slots_c, slots_f).items():
            self._buf(name, n, dtype)
        P = sc.max_hits
        scal = self._buf("scalars", 8, torch.int32)
        cs = self._stream("c_samples", slots_c)
        c_cnt = self._buf("c_cnt", R, torch.int32)
        check(L.nsvf_render_sample_coarse(
            _p(self._buf("hit_list", R, torch.int32)), _p(scal[0:1]), R,
            _p(self._buf("hits", R * P * 4, torch.int32)), _p(self._buf("count", R, torch.int32)),
            _p(self._buf("tot", R, torch.float32)), _p(self._buf("c_bound", R, torch.int32)),
            _p(self._buf("c_off", R, torch.int32)), P, sc.step_size, self.compat,
            self.fine_steps, _p(cs), _p(c_cnt), _p(self._buf("f_bound", R, torch.int32)),
            _p(scal[2:3]), *self._segment_args(), self._pass_seed(0), _stream(self.device)),
            "sample_coarse")
        self._mark("coarse_sample")
        return CoarseStream(R, n_hit, slots_c, self._buf("hits", R * P * 4, torch.int32),
                            self._buf("count", R, torch.int32), self._buf("tot", R, torch.float32),
                            self._buf("flag", R, torch.int32), self._buf("rank", R, torch.int32),
                            self._buf("hit_list", R, torch.int32),
                            self._buf("c_off", R, torch.int32), c_cnt, cs,
                            self._buf("f_bound", R, torch.int32), scal)

    @torch.no_grad()
    def coarse_stream(self, sc: DeviceScene, ray_start, ray_dir, order=None):
        """The geometry stages of a view: intersection, hit compaction and the
        coarse sample stream (encoder.py:498-556).  Returns a CoarseStream
        whose tensors are workspace views, valid until the next call.
        order: as in begin()."""
        self.begin(sc, ray_start, ray_dir, order=order)
        g = self._sample_coarse()
        self._pending = None
        return g

    # ---- phase 2: samples, field, compositing, background

    @torch.no_grad()
    def finish(self, keep=False):
        """Enqueues the rest of the view begun last (waits for its geometry
        if it is not done).  Returns (rgb [R,3], depth [R], missed [R])."""
        L = lib()
        dev = self.device
        st = _stream(dev)
        pd = self._pending
        sc, ray_start, ray_dir, R = pd["sc"], pd["rs"], pd["rd"], pd["R"]
        fs = self.fine_steps
        g = self._sample_coarse()
        n_hit, slots_c, scal, hit_list = g.n_hit, g.slots, g.scal, g.hit_list
        flag, rank, c_off, c_cnt, cs, f_bound = g.flag, g.rank, g.off, g.cnt, g.samples, g.f_bound
        slots_f = slots_c + n_hit * (int(math.ceil(fs)) + 2)
        c_field = self._buf("c_field", slots_c * 4, torch.float32)
        self._field(sc, ray_start, ray_dir, cs, scal[1:2], slots_c, c_field)
        if self.tolerance > 0:
            self._march(cs, slots_c, hit_list, n_hit, c_off, c_cnt, int(scal[2]), c_field, "c")
        self._mark("coarse_field")
        c_prob = self._buf("c_prob", slots_c, torch.float32)
        ray_rgb = self._buf("ray_rgb", R * 3, torch.float32)
        ray_depth = self._buf("ray_depth", R, torch.float32)
        ray_missed = self._buf("ray_missed", R, torch.float32)
        fine = self.hierarchical and n_hit > 0
        fold = fine and self.fold_composite
        if not fold:
            check(L.nsvf_render_composite(_p(hit_list), _p(scal[0:1]), max(n_hit, 1), _p(c_off),
                                          _p(c_cnt), _p(cs), _p(c_field), _p(c_prob),
                                          _p(ray_rgb), _p(ray_depth), _p(ray_missed), st),
                  "composite coarse")
            self._track(sc, hit_list, scal, n_hit, c_off, c_cnt, cs, c_prob, st)
        # in a hierarchical view the fine composite replaces the coarse colour
        # of every hit ray: folded, the fine sampler composites the coarse rows
        # itself (the same weights, into c_prob) and no coarse composite runs
        self._mark("coarse_composite")
        if fine:
            f_off = self._buf("f_off", R, torch.int32)
            self._scan(f_bound, f_off, R, scal[3:4])
            fsm = self._stream("f_samples", slots_f)
            f_cnt = self._buf("f_cnt", R, torch.int32)
            f_field = self._buf("f_field", slots_f * 4, torch.float32)
            reuse = self._reuse()
            src = None
            if reuse:
                src = self._buf("f_src", slots_f, torch.int16)
            check(L.nsvf_render_sample_fine(_p(hit_list), _p(scal[0:1]), R, _p(c_off), _p(c_cnt),
                                            _p(scal[2:3]), _p(cs), _p(c_field) if fold else None,
                                            _p(c_prob), _p(f_bound), _p(f_off), fs, self.compat,
                                            _p(fsm), _p(f_cnt), _p(src), *self._segment_args(),
                                            self._pass_seed(1), st), "sample_fine")
            if fold:
                self._track(sc, hit_list, scal, n_hit, c_off, c_cnt, cs, c_prob, st)
            if reuse:
                # the field's list of the fine slots to evaluate (src == -1),
                # its length in scal[4]
                nb = int(L.nsvf_render_eval_list_blocks(slots_f))
                bcnt = self._buf("e_bcnt", nb, torch.int32)
                boff = self._buf("e_boff", nb, torch.int32)
                e_list = self._buf("e_list", slots_f, torch.int32)
                check(L.nsvf_render_eval_list_count(_p(src), _p(scal[3:4]), slots_f, _p(bcnt), st),
                      "eval_list_count")
                self._scan(bcnt, boff, nb, scal[4:5])
                check(L.nsvf_render_eval_list(_p(src), _p(scal[3:4]), slots_f, _p(boff),
                                              _p(e_list), st), "eval_list")
            self._mark("fine_sample")
            if reuse:
                self._field(sc, ray_start, ray_dir, fsm, scal[4:5], slots_f, f_field, e_list)
            else:
                self._field(sc, ray_start, ray_dir, fsm, scal[3:4], slots_f, f_field)
            if self.tolerance > 0:
                S_f = int(f_cnt.index_select(0, hit_list[:n_hit].long()).max())
                self._march(fsm, slots_f, hit_list, n_hit, f_off, f_cnt, S_f, f_field, "f")
            self._mark("fine_field")
            f_prob = self._buf("f_prob", slots_f, torch.float32) if self.track is not None else None
            if reuse:
                check(L.nsvf_render_composite_reuse(
                    _p(hit_list), _p(scal[0:1]), n_hit, _p(f_off), _p(f_cnt), _p(fsm),
                    _p(f_field), _p(src), _p(c_off), _p(c_field), _p(f_prob), _p(ray_rgb),
                    _p(ray_depth), _p(ray_missed), st), "composite fine")
            else:
                check(L.nsvf_render_composite(_p(hit_list), _p(scal[0:1]), n_hit, _p(f_off),
                                              _p(f_cnt), _p(fsm), _p(f_field), _p(f_prob),
                                              _p(ray_rgb), _p(ray_depth), _p(ray_missed), st),
                      "composite fine")
            self._track(sc, hit_list, scal, n_hit, f_off, f_cnt, fsm, f_prob, st)
            self._mark("fine_composite")
        # the outputs leave the workspace: the caller owns them
        out = self._buf("out", R * 5, torch.float32)
        del self._ws["out"]
        rgb, depth, missed = out[:R * 3].view(R, 3), out[R * 3:R * 4], out[R * 4:R * 5]
        check(L.nsvf_render_finish(R, _p(flag), _p(ray_rgb), _p(ray_depth), _p(ray_missed),
                                   BG_COLOR[0], BG_COLOR[1], BG_COLOR[2], BG_DEPTH, _p(rgb),
                                   _p(depth), _p(missed), st), "render_finish")
        self._mark("finish")
        self._pending = None
        self.last = RenderStats(R, n_hit, slots_c, slots_f if self.hierarchical else 0, 0)
        if keep:
            hl = hit_list[:n_hit].long()
            P = sc.max_hits
            self.kept = {
                "hits": g.hits.view(R, P, 4), "count": g.count, "tot": g.tot, "flag": flag,
                "rank": rank, "hit_list": hit_list[:n_hit], "c_cnt": c_cnt[hl], "c_off": c_off[hl],
                "c_samples": _columns(cs), "c_field": c_field.view(-1, 4), "c_prob": c_prob,
                "c_max": int(scal[2]),
            }
            if self.tolerance > 0:
                self.kept.update({"c_eval": self._ws["c_eval"][:slots_c],
                                  "c_active": self._ws["c_active"][:n_hit]})
            if self.hierarchical and n_hit > 0:
                if reuse:   # the kept fine field holds every sample's value
                    check(L.nsvf_render_reuse_fill(_p(hit_list), _p(scal[0:1]), n_hit, _p(c_off),
                                                   _p(f_off), _p(f_cnt), _p(src), _p(c_field),
                                                   _p(f_field), st), "reuse_fill")
                self.kept.update({"f_cnt": f_cnt[hl], "f_off": f_off[hl],
                                  "f_samples": _columns(fsm), "f_field": f_field.view(-1, 4),
                                  "f_evaluated": int(scal[4]) if reuse else int(f_cnt[hl].sum())})
                if self.tolerance > 0:
                    self.kept.update({"f_eval": self._ws["f_eval"][:slots_f],
                                      "f_active": self._ws["f_active"][:n_hit]})
                if self.track is not None:
                    self.kept["f_prob"] = f_prob
        return rgb, depth, missed

    @torch.no_grad()
    def render_camera(self, sc: DeviceScene, camera, keep=False):
        """render() of the rays of a camera (nsvf_amd.rays.ViewCamera), which
        the traversal kernel generates; outputs in row-major pixel order."""
        self.begin_camera(sc, camera)
        return self.finish(keep)

    @torch.no_grad()
    def render(self, sc: DeviceScene, ray_start, ray_dir, keep=False, width=0):
        """ray_start, ray_dir: [R,3] float32 on the device.  Returns
        (rgb [R,3], depth [R], missed [R]) -- the fine pass when hierarchical,
        else the coarse pass -- after background fill.  width: as in begin()."""
        self.begin(sc, ray_start, ray_dir, width)
        return self.finish(keep)

    def count_samples(self):
        """Exact sample counts of the last view (reads the device)."""
        k = getattr(self, "kept", None)
        return None if k is None else (int(k["c_cnt"].sum()), int(k.get("f_cnt", k["c_cnt"]).sum()))


def _columns(rec):
    """(idx, depth, dist, ray) views of an int32 [m, 4] record buffer."""
    return (rec[:, 0], rec[:, 1].view(torch.float32), rec[:, 2].view(torch.float32), rec[:, 3])
