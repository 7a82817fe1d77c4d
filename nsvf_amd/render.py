"""Fast view renderer over the fused HIP pipeline, with the scene resident on
the device.

One call renders a batch of rays (a view, or several requests' rays) through
the stages of fairnr/models/fairnr_model.py:142-186 for the NSVF model with
hierarchical sampling:

  intersect (SVO) -> sort -> hit compaction -> coarse inverse-CDF samples ->
  field (gather + MLP) -> composite -> fine inverse-CDF samples over the
  coarse samples -> field -> composite -> background fill.

All stages are kernels of libnsvf_amd.so on the current stream.  Two device
to host reads per view size the sample streams (the coarse and fine sample
totals); the reference reads back far more (hit compaction, max lengths, a
per-column sum in its renderer loop).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import geometry
from ._lib import check, lib
from .field import pack_mlp
from .scene import VoxelScene

BG_DEPTH = 5.0              # fairnr/modules/implicit.py:19
BG_COLOR = (1.0, 1.0, 1.0)  # fairnr/modules/field.py:69 (transparent background)
PRECISIONS = {"bf16": 0, "fp32": 1}


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


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
        self.tree_centers = cen.float().contiguous().to(self.device)
        self.tree_children = ch.int().contiguous().to(self.device)
        self.feats = torch.from_numpy(scene.feats.astype(np.int32)).to(self.device)
        self.values = torch.from_numpy(scene.values).to(self.device)
        self.voxel_size = float(vs)
        self.step_size = float(np.float32(scene.step_size))
        self.max_hits = int(scene.max_hits)
        self.mlp = pack_mlp(scene.mlp, self.device)
        self.D, self.H, self.L = self.mlp["shape"]
        if not lib().nsvf_field_supported(self.D, self.H, self.L):
            raise RuntimeError(f"field shape D={self.D} H={self.H} L={self.L} not supported")

    @property
    def num_nodes(self):
        return self.tree_children.shape[0]

    def nbytes(self):
        ts = [self.centers, self.tree_centers, self.tree_children, self.feats, self.values,
              self.mlp["bf16"], self.mlp["f32"], self.mlp["bias"]]
        return sum(t.numel() * t.element_size() for t in ts)


@dataclass
class RenderStats:
    rays: int
    hit_rays: int
    coarse_samples: int
    fine_samples: int
    coarse_width: int


class Renderer:
    """Renders rays against a DeviceScene.  Workspace buffers are kept and
    grown as needed, so repeated views of the same size allocate nothing."""

    def __init__(self, device, precision="bf16", hierarchical=True, fine_steps=64.0,
                 compat=True):
        self.device = torch.device(device)
        self.precision = precision
        self.hierarchical = hierarchical
        self.fine_steps = float(fine_steps)
        self.compat = int(bool(compat))
        self._ws = {}
        self.last = None
        self.trace = None   # optional list of (stage, event) marks
        self._kernel_events = None  # optional list of (kernel, units, start, end)

    def record_kernels(self, on=True):
        """Start (or stop) recording events around the field-kernel launches;
        returns the list the records go to."""
        self._kernel_events = [] if on else None
        return self._kernel_events

    def _buf(self, name, n, dtype):
        t = self._ws.get(name)
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(max(int(n * 1.25), 1024), dtype=dtype, device=self.device)
            self._ws[name] = t
        return t[:n] if n else t[:0]

    def _mark(self, stage):
        if self.trace is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self.trace.append((stage, e))

    def _samples(self, prefix, m):
        return (self._buf(prefix + "idx", m, torch.int32), self._buf(prefix + "depth", m, torch.float32),
                self._buf(prefix + "dist", m, torch.float32), self._buf(prefix + "ray", m, torch.int32))

    def _field(self, sc, ray_start, ray_dir, samples, m_dev, m, out):
        L = lib()
        prec = PRECISIONS[self.precision]
        w = sc.mlp["bf16"] if prec == 0 else sc.mlp["f32"]
        ev = self._kernel_events is not None and m > 0
        if ev:
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record()
        check(L.nsvf_field_eval(prec, sc.D, sc.H, sc.L, _p(sc.centers), _p(sc.feats), _p(sc.values),
                                sc.voxel_size, _p(ray_start), _p(ray_dir), _p(samples[0]),
                                _p(samples[1]), _p(samples[3]), _p(m_dev), m, _p(w),
                                _p(sc.mlp["bias"]), _p(out), _stream(self.device)), "field_eval")
        if ev:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            self._kernel_events.append(("field", m, e0, e1))

    def _scan(self, counts, offsets, n, total):
        ws = self._buf("scan_ws", int(lib().nsvf_scan_workspace_ints(n)), torch.int32)
        check(lib().nsvf_scan_exclusive_i32(_p(counts), _p(offsets), n, _p(ws), _p(total),
                                            _stream(self.device)), "scan")

    @torch.no_grad()
    def render(self, sc: DeviceScene, ray_start, ray_dir, keep=False):
        """ray_start, ray_dir: [R,3] float32 on the device.  Returns
        (rgb [R,3], depth [R], missed [R]) -- the fine pass when hierarchical,
        else the coarse pass -- after background fill."""
        L = lib()
        dev = self.device
        st = _stream(dev)
        ray_start = ray_start.contiguous()
        ray_dir = ray_dir.contiguous()
        R = ray_start.shape[0]
        P = sc.max_hits
        self._mark("start")
        hits = self._buf("hits", R * P * 4, torch.int32)
        count = self._buf("count", R, torch.int32)
        tot = self._buf("tot", R, torch.float32)
        flag = self._buf("flag", R, torch.int32)
        check(L.nsvf_render_hits(_p(ray_start), _p(ray_dir), R, _p(sc.tree_centers),
                                 _p(sc.tree_children), sc.num_nodes, sc.voxel_size, P, _p(hits),
                                 _p(count), _p(tot), _p(flag), st), "render_hits")
        self._mark("intersect")
        scal = self._buf("scalars", 8, torch.int32)   # [n_hit, m_c, max_c, m_f, max_f]
        scal.zero_()
        rank = self._buf("rank", R, torch.int32)
        self._scan(flag, rank, R, scal[0:1])
        hit_list = self._buf("hit_list", R, torch.int32)
        check(L.nsvf_render_compact(_p(flag), _p(rank), R, _p(hit_list), st), "render_compact")
        # coarse samples
        c_cnt = self._buf("c_cnt", R, torch.int32)
        c_off = self._buf("c_off", R, torch.int32)
        check(L.nsvf_render_sample_coarse(0, _p(hit_list), _p(scal[0:1]), R, _p(hits), _p(count),
                                          _p(tot), P, sc.step_size, self.compat, _p(c_cnt),
                                          _p(scal[2:3]), None, None, None, None, None, st),
              "sample_coarse count")
        self._scan(c_cnt, c_off, R, scal[1:2])
        self._mark("coarse_count")
        host = scal[:3].cpu()
        n_hit, m_c, max_c = int(host[0]), int(host[1]), int(host[2])
        cs = self._samples("c_", m_c)
        check(L.nsvf_render_sample_coarse(1, _p(hit_list), _p(scal[0:1]), R, _p(hits), _p(count),
                                          _p(tot), P, sc.step_size, self.compat, None, None,
                                          _p(c_off), _p(cs[0]), _p(cs[1]), _p(cs[2]), _p(cs[3]), st),
              "sample_coarse write")
        self._mark("coarse_sample")
        c_field = self._buf("c_field", m_c * 4, torch.float32)
        self._field(sc, ray_start, ray_dir, cs, scal[1:2], m_c, c_field)
        self._mark("coarse_field")
        c_prob = self._buf("c_prob", m_c, torch.float32)
        ray_rgb = self._buf("ray_rgb", R * 3, torch.float32)
        ray_depth = self._buf("ray_depth", R, torch.float32)
        ray_missed = self._buf("ray_missed", R, torch.float32)
        check(L.nsvf_render_composite(_p(hit_list), _p(scal[0:1]), max(n_hit, 1), _p(c_off),
                                      _p(c_cnt), _p(cs[1]), _p(cs[2]), _p(c_field), _p(c_prob),
                                      _p(ray_rgb), _p(ray_depth), _p(ray_missed), st),
              "composite coarse")
        self._mark("coarse_composite")
        m_f = 0
        if self.hierarchical and n_hit > 0:
            f_cnt = self._buf("f_cnt", R, torch.int32)
            f_off = self._buf("f_off", R, torch.int32)
            check(L.nsvf_render_sample_fine(0, _p(hit_list), _p(scal[0:1]), R, _p(c_off), _p(c_cnt),
                                            _p(scal[2:3]), _p(cs[0]), _p(cs[1]), _p(cs[2]),
                                            _p(c_prob), self.fine_steps, self.compat, _p(f_cnt),
                                            _p(scal[4:5]), None, None, None, None, None, st),
                  "sample_fine count")
            self._scan(f_cnt, f_off, R, scal[3:4])
            m_f = int(scal[3:4].cpu())
            fs = self._samples("f_", m_f)
            check(L.nsvf_render_sample_fine(1, _p(hit_list), _p(scal[0:1]), R, _p(c_off), _p(c_cnt),
                                            _p(scal[2:3]), _p(cs[0]), _p(cs[1]), _p(cs[2]),
                                            _p(c_prob), self.fine_steps, self.compat, None, None,
                                            _p(f_off), _p(fs[0]), _p(fs[1]), _p(fs[2]), _p(fs[3]),
                                            st), "sample_fine write")
            self._mark("fine_sample")
            f_field = self._buf("f_field", m_f * 4, torch.float32)
            self._field(sc, ray_start, ray_dir, fs, scal[3:4], m_f, f_field)
            self._mark("fine_field")
            check(L.nsvf_render_composite(_p(hit_list), _p(scal[0:1]), n_hit, _p(f_off), _p(f_cnt),
                                          _p(fs[1]), _p(fs[2]), _p(f_field), None, _p(ray_rgb),
                                          _p(ray_depth), _p(ray_missed), st), "composite fine")
            self._mark("fine_composite")
        rgb = torch.empty((R, 3), dtype=torch.float32, device=dev)
        depth = torch.empty(R, dtype=torch.float32, device=dev)
        missed = torch.empty(R, dtype=torch.float32, device=dev)
        check(L.nsvf_render_finish(R, _p(flag), _p(rank), _p(ray_rgb), _p(ray_depth), _p(ray_missed),
                                   BG_COLOR[0], BG_COLOR[1], BG_COLOR[2], BG_DEPTH, _p(rgb),
                                   _p(depth), _p(missed), st), "render_finish")
        self._mark("finish")
        self.last = RenderStats(R, n_hit, m_c, m_f, max_c)
        if keep:
            self.kept = {
                "hits": hits.view(R, P, 4), "count": count, "tot": tot, "flag": flag, "rank": rank,
                "hit_list": hit_list[:n_hit], "c_cnt": c_cnt[:n_hit], "c_off": c_off[:n_hit],
                "c_samples": cs, "c_field": c_field.view(-1, 4), "c_prob": c_prob,
            }
            if m_f:
                self.kept.update({"f_cnt": f_cnt[:n_hit], "f_off": f_off[:n_hit],
                                  "f_samples": fs, "f_field": f_field.view(-1, 4)})
        return rgb, depth, missed
