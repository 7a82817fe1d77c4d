"""The reference's Python op forms, ``fairnr.clib`` (fairnr/clib/__init__.py),
as a drop-in: ``aabb_ray_intersect``, ``svo_ray_intersect``,
``uniform_ray_sampling``, ``inverse_cdf_sampling`` with the same call
signatures, autograd behaviour (outputs non-differentiable, backward returns
``None``s) and outputs.

What changes underneath:
  * no ``expand(...).contiguous()`` copies of the scene per ray group
    (clib/__init__.py:71,110-111): the kernels read one shared copy;
  * the intersection ops run once over all rays instead of G regrouped
    blocks (each ray's result is independent of the grouping);
  * ``inverse_cdf_sampling`` runs one launch over all rows instead of a
    launch per 800-column chunk plus ``torch.cat`` (clib/__init__.py:259-274);
    the kernel derives each ray's neighbours under the reference's G=200
    regrouping, which the trailing-loop quirks read, so the output is the
    reference's;
  * with ``deterministic=True`` no noise tensor is allocated (the kernel
    uses the constant 0.5).
"""
from __future__ import annotations

import math

import torch
from torch.autograd import Function

from . import _ext
from ._ext import _ptr, _stream, _scalar
from ._lib import check, lib

MAX_DEPTH = 10000.0   # fairnr/clib/__init__.py:36

__all__ = ["aabb_ray_intersect", "svo_ray_intersect", "uniform_ray_sampling",
           "inverse_cdf_sampling", "MAX_DEPTH"]


class AABBRayIntersect(Function):
    """clib/__init__.py:58-95."""

    @staticmethod
    def forward(ctx, voxelsize, n_max, points, ray_start, ray_dir):
        pts = points.float().contiguous()
        if pts.dim() == 2:
            pts = pts.unsqueeze(0)
        inds, min_depth, max_depth = _ext.aabb_intersect(
            ray_start.float().contiguous(), ray_dir.float().contiguous(), pts, voxelsize, n_max)
        min_depth = min_depth.type_as(ray_start)
        max_depth = max_depth.type_as(ray_start)
        ctx.mark_non_differentiable(inds)
        ctx.mark_non_differentiable(min_depth)
        ctx.mark_non_differentiable(max_depth)
        return inds, min_depth, max_depth

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None


aabb_ray_intersect = AABBRayIntersect.apply


class SparseVoxelOctreeRayIntersect(Function):
    """clib/__init__.py:98-135."""

    @staticmethod
    def forward(ctx, voxelsize, n_max, points, children, ray_start, ray_dir):
        pts = points.float().contiguous()
        ch = children.int().contiguous()
        if pts.dim() == 2:
            pts, ch = pts.unsqueeze(0), ch.unsqueeze(0)
        inds, min_depth, max_depth = _ext.svo_intersect(
            ray_start.float().contiguous(), ray_dir.float().contiguous(), pts, ch, voxelsize, n_max)
        min_depth = min_depth.type_as(ray_start)
        max_depth = max_depth.type_as(ray_start)
        ctx.mark_non_differentiable(inds)
        ctx.mark_non_differentiable(min_depth)
        ctx.mark_non_differentiable(max_depth)
        return inds, min_depth, max_depth

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None


svo_ray_intersect = SparseVoxelOctreeRayIntersect.apply


def _trim(idx, depth, dists):
    max_len = int(idx.ne(-1).sum(-1).max()) if idx.numel() else 0
    return idx[:, :max_len], depth[:, :max_len], dists[:, :max_len]


class UniformRaySampling(Function):
    """clib/__init__.py:178-228.  The G=256 regrouping is contiguous, so every
    ray reads the same neighbours in one [1,N,P] launch."""

    @staticmethod
    def forward(ctx, pts_idx, min_depth, max_depth, step_size, max_ray_length, deterministic=False):
        N, P = pts_idx.size(0), pts_idx.size(1)
        max_steps = int(_scalar(max_ray_length) / _scalar(step_size))
        max_steps = max_steps + P * 2
        noise = min_depth.new_zeros(1, N, max_steps, dtype=torch.float32)
        if deterministic:
            noise += 0.5
        else:
            noise = noise.uniform_()
        si, sd, sx = _ext.uniform_ray_sampling(
            pts_idx.int().contiguous().unsqueeze(0), min_depth.float().contiguous().unsqueeze(0),
            max_depth.float().contiguous().unsqueeze(0), noise, step_size, max_steps)
        si, sd, sx = si[0], sd[0].type_as(min_depth), sx[0].type_as(min_depth)
        si, sd, sx = _trim(si, sd, sx)
        ctx.mark_non_differentiable(si)
        ctx.mark_non_differentiable(sd)
        ctx.mark_non_differentiable(sx)
        return si, sd, sx

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None, None


uniform_ray_sampling = UniformRaySampling.apply

ICDF_GROUPS = 200        # G, clib/__init__.py:234
ICDF_CHUNK = 4 * ICDF_GROUPS  # columns per launch in the reference, :259


def inverse_cdf_sampling_rows(pts_idx, min_depth, max_depth, probs, steps, fixed_step_size=-1.0,
                              noise=None, max_steps=None, compat=True):
    """One launch over [N,P] rows with the reference's grouping semantics.
    Returns untrimmed [N, max_steps] outputs and max_steps."""
    N, P = pts_idx.shape
    dev = pts_idx.device
    if max_steps is None:
        max_steps = int(steps.ceil().max().item()) + P if N > 0 else P
    pi = pts_idx.int().contiguous()
    mn, mx = min_depth.float().contiguous(), max_depth.float().contiguous()
    pr, st = probs.float().contiguous(), steps.float().contiguous()
    if noise is not None:
        noise = noise.float().contiguous()
        if tuple(noise.shape) != (N, max_steps):
            raise RuntimeError("noise must be [N, max_steps]")
    oi = torch.empty((N, max_steps), dtype=torch.int32, device=dev)
    od = torch.empty((N, max_steps), dtype=torch.float32, device=dev)
    ot = torch.empty((N, max_steps), dtype=torch.float32, device=dev)
    check(lib().nsvf_inverse_cdf_sampling_rows(_ptr(pi), _ptr(mn), _ptr(mx), _ptr(noise), _ptr(pr),
                                               _ptr(st), N, P, max_steps, ICDF_GROUPS, ICDF_CHUNK,
                                               float(_scalar(fixed_step_size)), int(bool(compat)),
                                               _ptr(oi), _ptr(od), _ptr(ot), None, _stream(dev)),
          "inverse_cdf_sampling")
    return oi, od, ot, max_steps


class InverseCDFRaySampling(Function):
    """clib/__init__.py:231-300."""

    @staticmethod
    def forward(ctx, pts_idx, min_depth, max_depth, probs, steps, fixed_step_size=-1,
                deterministic=False):
        N, P = pts_idx.size(0), pts_idx.size(1)
        max_steps = int(steps.ceil().max().item()) + P if N > 0 else P
        noise = None
        if not deterministic:
            noise = min_depth.new_zeros(N, max_steps, dtype=torch.float32).uniform_().clamp(
                min=0.001, max=0.999)
        si, sd, sx, _ = inverse_cdf_sampling_rows(pts_idx, min_depth, max_depth, probs, steps,
                                                  fixed_step_size, noise, max_steps)
        sd, sx = sd.type_as(min_depth), sx.type_as(min_depth)
        si, sd, sx = _trim(si, sd, sx)
        ctx.mark_non_differentiable(si)
        ctx.mark_non_differentiable(sd)
        ctx.mark_non_differentiable(sx)
        return si, sd, sx

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None, None, None


inverse_cdf_sampling = InverseCDFRaySampling.apply
