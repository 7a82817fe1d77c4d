"""The plain FP32 GPU path: the reference's NSVF view rendering in its naive
form on MI355X, the comparator the fused renderer is measured against.

It follows the reference's Python orchestration step by step, with torch ops
where the reference uses torch ops and the naive comparator kernels (one
thread per ray, reference launch geometry, no LDS) where it calls
`fairnr.clib._ext`:

  fairnr_model.py:142-186   _forward: precompute, intersect, hit compaction,
                            raymarching, hierarchical pass, postprocessing
  encoder.py:380-410        precompute (x-shift of the centres)
  encoder.py:498-535        ray_intersect (+ sort, pads to MAX_DEPTH)
  clib/__init__.py:98-135   svo wrapper: G groups of K rays, expanded copies
                            of the octree per group
  nsvf.py:65-74             probs and steps
  clib/__init__.py:231-300  inverse-CDF wrapper: G=200 regrouping, one launch
                            per 800 columns, torch.cat, trim to max length
  encoder.py:538-592        ray_sample tail, trilinear gather (F.embedding)
  renderer.py:77-252        forward_chunk column loop with a host sync per
                            column, forward_once per <= 65,536 samples, the
                            field as FP32 nn.Linear layers, compositing
  nerf.py:64-79             hierarchical resampling over the coarse samples
  nsvf.py:89-101            fill-in and background
"""
from __future__ import annotations

import ctypes
import math
from collections import defaultdict

import numpy as np
import torch
import torch.nn.functional as F

from . import geometry
from ._lib import check, lib
from .scene import VoxelScene

MAX_DEPTH = 10000.0
BG_DEPTH = 5.0


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


# Geometry helpers of the reference's torch orchestration
# (fairnr/data/geometry.py), used only by this comparator.

def offset_points(point_xyz, quarter_voxel=1.0, offset_only=False, bits=2):
    """geometry.py:229-238.  bits^3 lattice offsets (2i+1-bits)/(bits-1) *
    quarter_voxel around each point, x slowest and z fastest (meshgrid 'ij');
    with bits=2 these are the 8 corners in the order 4x + 2y + z."""
    c = torch.arange(1, 2 * bits, 2, device=point_xyz.device)
    ox, oy, oz = torch.meshgrid([c, c, c], indexing="ij")
    offset = (torch.cat([ox.reshape(-1, 1), oy.reshape(-1, 1), oz.reshape(-1, 1)], 1)
              .type_as(point_xyz) - bits) / float(bits - 1)
    if not offset_only:
        return point_xyz.unsqueeze(1) + offset.unsqueeze(0).type_as(point_xyz) * quarter_voxel
    return offset.type_as(point_xyz) * quarter_voxel


def trilinear_interp(p, q, point_feats):
    """geometry.py:195-200: weights prod(p*q + (1-p)(1-q)) over the 8 corners."""
    weights = (p * q + (1 - p) * (1 - q)).prod(dim=-1, keepdim=True)
    if point_feats.dim() == 2:
        point_feats = point_feats.view(point_feats.size(0), 8, -1)
    return (weights * point_feats).sum(1)


def fill_in(shape, hits, input, initial=1.0):
    """geometry.py:303-317: scatter the hit rays' values back to the frame."""
    input_sizes = list(input.size())
    if len(input_sizes) == len(shape) and all(shape[i] == input_sizes[i] for i in range(len(shape))):
        return input
    if isinstance(initial, torch.Tensor):
        output = initial.expand(*shape)
    else:
        output = input.new_ones(*shape) * initial
    if len(shape) == 1:
        return output.masked_scatter(hits, input)
    return output.masked_scatter(hits.unsqueeze(-1).expand(*shape), input)


def _svo_ray_intersect(voxelsize, n_max, points, children, ray_start, ray_dir):
    """clib/__init__.py:98-135 with the naive kernel."""
    G = min(2048, int(2 * 10 ** 9 / (points.numel() + children.numel())))
    S, N = ray_start.shape[:2]
    K = int(np.ceil(N / G))
    H = K * G
    if H > N:
        ray_start = torch.cat([ray_start, ray_start[:, :H - N]], 1)
        ray_dir = torch.cat([ray_dir, ray_dir[:, :H - N]], 1)
    ray_start = ray_start.reshape(S * G, K, 3).contiguous()
    ray_dir = ray_dir.reshape(S * G, K, 3).contiguous()
    points = points.expand(S * G, *points.size()[1:]).contiguous()
    children = children.expand(S * G, *children.size()[1:]).contiguous()
    B, M, T = S * G, K, points.size(1)
    idx = torch.zeros((B, M, n_max), dtype=torch.int32, device=ray_start.device)
    lo = torch.zeros((B, M, n_max), dtype=torch.float32, device=ray_start.device)
    hi = torch.zeros((B, M, n_max), dtype=torch.float32, device=ray_start.device)
    check(lib().nsvf_plain_svo_intersect(_p(ray_start), _p(ray_dir), _p(points), _p(children), B, M,
                                         T, float(voxelsize), int(n_max), _p(idx), _p(lo), _p(hi),
                                         _stream(ray_start.device)), "plain_svo_intersect")
    idx, lo, hi = idx.reshape(S, H, -1), lo.reshape(S, H, -1), hi.reshape(S, H, -1)
    if H > N:
        idx, lo, hi = idx[:, :N], lo[:, :N], hi[:, :N]
    return idx, lo, hi


def _inverse_cdf_sampling(pts_idx, min_depth, max_depth, probs, steps, fixed_step_size=-1,
                          deterministic=True, compat=True):
    """clib/__init__.py:231-300 with the naive kernel."""
    G, N, P = 200, pts_idx.size(0), pts_idx.size(1)
    H = int(np.ceil(N / G)) * G
    if H > N:
        pts_idx = torch.cat([pts_idx, pts_idx[:1].expand(H - N, P)], 0)
        min_depth = torch.cat([min_depth, min_depth[:1].expand(H - N, P)], 0)
        max_depth = torch.cat([max_depth, max_depth[:1].expand(H - N, P)], 0)
        probs = torch.cat([probs, probs[:1].expand(H - N, P)], 0)
        steps = torch.cat([steps, steps[:1].expand(H - N)], 0)
    pts_idx = pts_idx.reshape(G, -1, P)
    min_depth = min_depth.reshape(G, -1, P)
    max_depth = max_depth.reshape(G, -1, P)
    probs = probs.reshape(G, -1, P)
    steps = steps.reshape(G, -1)
    max_steps = int(steps.ceil().long().max()) + P
    noise = min_depth.new_zeros(*min_depth.size()[:-1], max_steps)
    if deterministic:
        noise += 0.5
    else:
        noise = noise.uniform_().clamp(min=0.001, max=0.999)
    chunk_size = 4 * G
    results = []
    for i in range(0, min_depth.size(1), chunk_size):
        a = [pts_idx[:, i:i + chunk_size].int().contiguous(),
             min_depth.float()[:, i:i + chunk_size].contiguous(),
             max_depth.float()[:, i:i + chunk_size].contiguous(),
             noise.float()[:, i:i + chunk_size].contiguous(),
             probs.float()[:, i:i + chunk_size].contiguous(),
             steps.float()[:, i:i + chunk_size].contiguous()]
        B, R = a[0].shape[:2]
        oi = -torch.ones((B, R, max_steps), dtype=torch.int32, device=a[0].device)
        od = torch.zeros((B, R, max_steps), dtype=torch.float32, device=a[0].device)
        ox = torch.zeros((B, R, max_steps), dtype=torch.float32, device=a[0].device)
        check(lib().nsvf_plain_inverse_cdf_sampling(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), _p(a[4]),
                                                    _p(a[5]), B, R, P, max_steps,
                                                    float(fixed_step_size), int(bool(compat)),
                                                    _p(oi), _p(od), _p(ox),
                                                    _stream(a[0].device)), "plain_icdf")
        results.append((oi, od, ox))
    si, sd, sx = [torch.cat([r[i] for r in results], 1) for i in range(3)]
    si, sd, sx = si.reshape(H, -1), sd.reshape(H, -1), sx.reshape(H, -1)
    if H > N:
        si, sd, sx = si[:N], sd[:N], sx[:N]
    max_len = si.ne(-1).sum(-1).max()
    return si[:, :max_len], sd[:, :max_len], sx[:, :max_len]


class PlainNSVF:
    """The reference NSVF model's eval-time rendering for one scene."""

    def __init__(self, scene: VoxelScene, device, chunk_size=64, hierarchical=True,
                 fine_num_samples=64, compat=True):
        dev = torch.device(device)
        self.device = dev
        self.points = torch.from_numpy(scene.points).to(dev)
        self.feats = torch.from_numpy(scene.feats).long().to(dev)
        self.keep = torch.ones(self.points.size(0), dtype=torch.long, device=dev)
        self.values = torch.from_numpy(scene.values).to(dev)
        self.voxel_size = torch.scalar_tensor(scene.voxel_size).to(dev)
        self.step_size = torch.scalar_tensor(scene.step_size).to(dev)
        self.max_hits = torch.scalar_tensor(scene.max_hits).to(dev)
        self.layers = [(torch.from_numpy(w).to(dev), torch.from_numpy(b).to(dev))
                       for w, b in zip(scene.mlp.weights, scene.mlp.biases)]
        self.chunk_size = 1024 * chunk_size
        self.hierarchical = hierarchical
        self.fine_num_samples = fine_num_samples
        # compat=False: the sampler's evident intent (no trailing duplicate
        # sample carrying the next ray's voxel id, sample_gpu.cu:188)
        self.compat = compat
        self._octree = None

    # encoder.py:365-378 (runtime cache, built lazily from the kept points)
    def octree(self):
        if self._octree is None:
            pts = self.points[self.keep.bool()]
            self._octree = geometry.build_easy_octree(pts, self.voxel_size / 2.0)
        return self._octree

    def precompute(self):
        feats = self.feats[self.keep.bool()]
        points = self.points[self.keep.bool()]
        points[:, 0] += (self.voxel_size / 10)
        cen, ch = self.octree()
        return {"voxel_vertex_idx": feats, "voxel_center_xyz": points,
                "voxel_vertex_emb": self.values,
                "voxel_octree_center_xyz": cen.clone(), "voxel_octree_children_idx": ch.clone()}

    def ray_intersect(self, ray_start, ray_dir, st):
        S = 1
        ray_start = ray_start.reshape(S, -1, 3).contiguous()
        ray_dir = ray_dir.reshape(S, -1, 3).contiguous()
        pts_idx, min_depth, max_depth = _svo_ray_intersect(
            self.voxel_size.item(), int(self.max_hits.item()), st["voxel_octree_center_xyz"][None],
            st["voxel_octree_children_idx"][None].int(), ray_start, ray_dir)
        min_depth.masked_fill_(pts_idx.eq(-1), MAX_DEPTH)
        max_depth.masked_fill_(pts_idx.eq(-1), MAX_DEPTH)
        min_depth, sorted_idx = min_depth.sort(dim=-1)
        max_depth = max_depth.gather(-1, sorted_idx)
        pts_idx = pts_idx.gather(-1, sorted_idx)
        hits = pts_idx.ne(-1).any(-1)
        return {"min_depth": min_depth, "max_depth": max_depth, "intersected_voxel_idx": pts_idx}, hits

    def ray_sample(self, io):
        si, sd, sx = _inverse_cdf_sampling(io["intersected_voxel_idx"], io["min_depth"],
                                           io["max_depth"], io["probs"], io["steps"], -1, True,
                                           self.compat)
        sx = sx.clamp(min=0.0)
        sd.masked_fill_(si.eq(-1), MAX_DEPTH)
        sx.masked_fill_(si.eq(-1), 0.0)
        return {"sampled_point_depth": sd, "sampled_point_distance": sx,
                "sampled_point_voxel_idx": si}

    def encoder_forward(self, samples, st):
        idx = samples["sampled_point_voxel_idx"].long()
        xyz = samples["sampled_point_xyz"]
        point_xyz = F.embedding(idx, st["voxel_center_xyz"])
        point_feats = F.embedding(F.embedding(idx, st["voxel_vertex_idx"]),
                                  st["voxel_vertex_emb"]).view(point_xyz.size(0), -1)
        p = ((xyz - point_xyz) / self.voxel_size + .5).unsqueeze(1)
        q = offset_points(p, .5, offset_only=True).unsqueeze(0) + .5
        return {"emb": trilinear_interp(p, q, point_feats),
                "dists": samples["sampled_point_distance"]}

    def field(self, emb):
        h = emb
        for i, (w, b) in enumerate(self.layers):
            h = F.linear(h, w, b)
            if i < len(self.layers) - 1:
                h = torch.relu(h)
        return h[:, 0], torch.sigmoid(h[:, 1:4])

    def forward_once(self, ray_start, ray_dir, samples, st):
        sampled_depth = samples["sampled_point_depth"]
        sampled_idx = samples["sampled_point_voxel_idx"].long()
        sample_mask = sampled_idx.ne(-1)
        if sample_mask.sum() == 0:
            return None, 0
        samples = dict(samples)
        samples["sampled_point_xyz"] = ray_start.unsqueeze(1) + ray_dir.unsqueeze(1) * sampled_depth.unsqueeze(2)
        samples = {k: s[sample_mask] for k, s in samples.items()}
        inputs = self.encoder_forward(samples, st)
        sigma, tex = self.field(inputs["emb"])
        free_energy = torch.relu(sigma) * inputs["dists"] * 7.0
        B, K = sample_mask.size()
        out = {"sample_mask": sample_mask,
               "free_energy": free_energy.new_zeros(B, K).masked_scatter(sample_mask, free_energy),
               "texture": tex.new_zeros(B, K, 3).masked_scatter(
                   sample_mask.unsqueeze(-1).expand(B, K, 3), tex)}
        return out, sample_mask.sum()

    def forward_chunk(self, ray_start, ray_dir, samples, st):
        sampled_depth = samples["sampled_point_depth"]
        sampled_idx = samples["sampled_point_voxel_idx"].long()
        hits = sampled_idx.ne(-1).long()
        outputs = defaultdict(list)
        size_so_far, start_step = 0, 0
        for i in range(hits.size(1) + 1):
            if ((i == hits.size(1)) or (size_so_far + hits[:, i].sum() > self.chunk_size)) and (i > start_step):
                _o, _ = self.forward_once(ray_start, ray_dir,
                                          {k: s[:, start_step:i] for k, s in samples.items()}, st)
                if _o is not None:
                    for k in _o:
                        outputs[k].append(_o[k])
                else:
                    for k in outputs:
                        outputs[k].append(outputs[k][-1].new_zeros(
                            outputs[k][-1].size(0), sampled_depth[:, start_step:i].size(1),
                            *outputs[k][-1].size()[2:]))
                start_step, size_so_far = i, 0
            if i < hits.size(1):
                size_so_far += hits[:, i].sum()
        outputs = {k: torch.cat(outputs[k], 1) for k in outputs}
        fe = outputs["free_energy"]
        shifted = torch.cat([fe.new_zeros(sampled_depth.size(0), 1), fe[:, :-1]], dim=-1)
        a = 1 - torch.exp(-fe.float())
        b = torch.exp(-torch.cumsum(shifted.float(), dim=-1))
        probs = (a * b).type_as(fe)
        return {"probs": probs, "depths": (sampled_depth * probs).sum(-1),
                "missed": 1 - probs.sum(-1),
                "colors": (outputs["texture"] * probs.unsqueeze(-1)).sum(-2)}

    def raymarch(self, ray_start, ray_dir, io, st):
        samples = self.ray_sample(io)
        if ray_start.size(0) <= self.chunk_size:
            res = self.forward_chunk(ray_start, ray_dir, samples, st)
        else:
            parts = [self.forward_chunk(ray_start[i:i + self.chunk_size], ray_dir[i:i + self.chunk_size],
                                        {k: s[i:i + self.chunk_size] for k, s in samples.items()}, st)
                     for i in range(0, ray_start.size(0), self.chunk_size)]
            res = {k: torch.cat([p[k] for p in parts], 0) for k in parts[0]}
        return samples, res

    @torch.no_grad()
    def render(self, ray_start, ray_dir):
        """ray_start, ray_dir: [R,3].  Returns (rgb [R,3], depth [R], missed [R])."""
        return self.forward(ray_start, ray_dir)

    def forward(self, ray_start, ray_dir):
        """render() with autograd: differentiable in self.values and the
        tensors of self.layers when they require grad (the training forward,
        rendering_loss.py:37 -> fairnr_model.py:142-186, with eval-mode
        sampling noise 0.5)."""
        st = self.precompute()
        R = ray_dir.shape[0]
        io, hits = self.ray_intersect(ray_start, ray_dir, st)
        hits = hits.reshape(-1)
        rgb = torch.zeros((R, 3), device=self.device)
        depth = torch.zeros(R, device=self.device)
        missed = torch.ones(R, device=self.device)
        if hits.sum() > 0:
            io = {k: v[0][hits] for k, v in io.items()}
            rs, rd = ray_start[hits], ray_dir[hits]
            dists = (io["max_depth"] - io["min_depth"]).masked_fill(io["intersected_voxel_idx"].eq(-1), 0)
            io["probs"] = dists / dists.sum(dim=-1, keepdim=True)
            io["steps"] = dists.sum(-1) / self.step_size
            samples, res = self.raymarch(rs, rd, io, st)
            if self.hierarchical:
                depth_c = samples["sampled_point_depth"]
                dists_c = samples["sampled_point_distance"]
                io = {"min_depth": depth_c - dists_c * .5, "max_depth": depth_c + dists_c * .5,
                      "intersected_voxel_idx": samples["sampled_point_voxel_idx"].contiguous()}
                safe = res["probs"] + 1e-5
                io["probs"] = safe / safe.sum(-1, keepdim=True)
                io["steps"] = safe.new_ones(*safe.size()[:-1]) * self.fine_num_samples
                samples, res = self.raymarch(rs, rd, io, st)
            rgb = fill_in((R, 3), hits, res["colors"], 0.0)
            depth = fill_in((R,), hits, res["depths"], 0.0)
            missed = fill_in((R,), hits, res["missed"], 1.0)
        rgb = rgb + missed.unsqueeze(-1) * torch.ones(3, device=self.device)
        depth = depth + missed * BG_DEPTH
        return rgb, depth, missed
