"""The reference's native op module, ``fairnr.clib._ext``
(fairnr/clib/src/binding.cpp:11-21), over the C ABI of ``libnsvf_amd.so``.

Same names, argument order, dtypes and padded outputs.  Argument checks raise
``RuntimeError`` as the reference's ``TORCH_CHECK`` does (fairnr/clib/include/
utils.h:10-30); a HIP failure raises too, where the reference calls exit(-1)
(fairnr/clib/include/cuda_utils.h:35-44).  Kernels run asynchronously on the
current stream.  ``points``/``children`` may be given with batch 1 to share one
scene across all ray batches (the reference's callers pass expanded copies;
both forms are accepted).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import NSVFError, check, lib

__all__ = [
    "aabb_intersect", "svo_intersect", "uniform_ray_sampling", "inverse_cdf_sampling",
    "build_octree", "ball_intersect", "triangle_intersect",
]


def _check_cuda(name, t):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise RuntimeError(f"{name} must be a CUDA tensor")


def _check_contig(name, t):
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be a contiguous tensor")


def _check_float(name, t):
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float tensor")


def _check_int(name, t):
    if t.dtype != torch.int32:
        raise RuntimeError(f"{name} must be an int tensor")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _scalar(x):
    return x.item() if isinstance(x, torch.Tensor) else x


def _same_device(*ts):
    d = ts[0].device
    for t in ts[1:]:
        if t.device != d:
            raise RuntimeError("all tensors must be on the same device")


def aabb_intersect(ray_start, ray_dir, points, voxelsize, n_max):
    """intersect.cpp:49-75.  ray_start, ray_dir [B,M,3]; points [B,N,3] or
    [1,N,3].  Returns (idx i32, min_depth f32, max_depth f32), [B,M,n_max]."""
    for n, t in (("ray_start", ray_start), ("ray_dir", ray_dir), ("points", points)):
        _check_contig(n, t)
        _check_float(n, t)
        _check_cuda(n, t)
    _same_device(ray_start, ray_dir, points)
    if ray_start.dim() != 3 or ray_start.shape != ray_dir.shape or ray_start.size(2) != 3:
        raise RuntimeError("ray_start and ray_dir must be [B,M,3]")
    if points.dim() != 3 or points.size(2) != 3 or points.size(0) not in (1, ray_start.size(0)):
        raise RuntimeError("points must be [B,N,3] or [1,N,3]")
    n_max = int(_scalar(n_max))
    B, M = ray_start.shape[:2]
    dev = ray_start.device
    idx = torch.empty((B, M, n_max), dtype=torch.int32, device=dev)
    lo = torch.empty((B, M, n_max), dtype=torch.float32, device=dev)
    hi = torch.empty((B, M, n_max), dtype=torch.float32, device=dev)
    shared = int(points.size(0) == 1)
    check(lib().nsvf_aabb_intersect(_ptr(ray_start), _ptr(ray_dir), _ptr(points), B, M,
                                    points.size(1), shared, float(_scalar(voxelsize)), n_max,
                                    _ptr(idx), _ptr(lo), _ptr(hi), _stream(dev)), "aabb_intersect")
    return idx, lo, hi


def svo_intersect(ray_start, ray_dir, points, children, voxelsize, n_max):
    """intersect.cpp:84-112.  points [B,T,3] f32 (or [1,T,3]); children
    [B,T,9] i32.  Returns (idx, min_depth, max_depth), [B,M,n_max]."""
    for n, t in (("ray_start", ray_start), ("ray_dir", ray_dir), ("points", points),
                 ("children", children)):
        _check_contig(n, t)
        _check_cuda(n, t)
    for n, t in (("ray_start", ray_start), ("ray_dir", ray_dir), ("points", points)):
        _check_float(n, t)
    _check_int("children", children)
    _same_device(ray_start, ray_dir, points, children)
    if ray_start.dim() != 3 or ray_start.shape != ray_dir.shape or ray_start.size(2) != 3:
        raise RuntimeError("ray_start and ray_dir must be [B,M,3]")
    if (points.dim() != 3 or points.size(2) != 3 or children.dim() != 3 or children.size(2) != 9
            or children.shape[:2] != points.shape[:2]
            or points.size(0) not in (1, ray_start.size(0))):
        raise RuntimeError("points must be [B,T,3] and children [B,T,9] (B = 1 or ray batches)")
    n_max = int(_scalar(n_max))
    B, M = ray_start.shape[:2]
    dev = ray_start.device
    idx = torch.empty((B, M, n_max), dtype=torch.int32, device=dev)
    lo = torch.empty((B, M, n_max), dtype=torch.float32, device=dev)
    hi = torch.empty((B, M, n_max), dtype=torch.float32, device=dev)
    shared = int(points.size(0) == 1)
    check(lib().nsvf_svo_intersect(_ptr(ray_start), _ptr(ray_dir), _ptr(points), _ptr(children),
                                   B, M, points.size(1), shared, float(_scalar(voxelsize)), n_max,
                                   _ptr(idx), _ptr(lo), _ptr(hi), _stream(dev)), "svo_intersect")
    return idx, lo, hi


def uniform_ray_sampling(pts_idx, min_depth, max_depth, uniform_noise, step_size, max_steps):
    """sample.cpp:23-55.  pts_idx [B,R,P] i32, depths [B,R,P] f32, noise
    [B,R,max_steps] f32.  Returns (idx, depth, dists), [B,R,max_steps]."""
    for n, t in (("pts_idx", pts_idx), ("min_depth", min_depth), ("max_depth", max_depth),
                 ("uniform_noise", uniform_noise)):
        _check_contig(n, t)
        _check_cuda(n, t)
    for n, t in (("min_depth", min_depth), ("max_depth", max_depth),
                 ("uniform_noise", uniform_noise)):
        _check_float(n, t)
    _check_int("pts_idx", pts_idx)
    _same_device(pts_idx, min_depth, max_depth, uniform_noise)
    B, R, P = pts_idx.shape
    S = int(_scalar(max_steps))
    if min_depth.shape != pts_idx.shape or max_depth.shape != pts_idx.shape or \
            tuple(uniform_noise.shape) != (B, R, S):
        raise RuntimeError("shape mismatch: pts_idx/min/max [B,R,P], noise [B,R,max_steps]")
    dev = pts_idx.device
    oi = torch.empty((B, R, S), dtype=torch.int32, device=dev)
    od = torch.empty((B, R, S), dtype=torch.float32, device=dev)
    ot = torch.empty((B, R, S), dtype=torch.float32, device=dev)
    check(lib().nsvf_uniform_ray_sampling(_ptr(pts_idx), _ptr(min_depth), _ptr(max_depth),
                                          _ptr(uniform_noise), B, R, P, S,
                                          float(_scalar(step_size)), _ptr(oi), _ptr(od), _ptr(ot),
                                          None, _stream(dev)), "uniform_ray_sampling")
    return oi, od, ot


def inverse_cdf_sampling(pts_idx, min_depth, max_depth, uniform_noise, probs, steps,
                         fixed_step_size, compat=True):
    """sample.cpp:58-95.  pts_idx [B,R,P] i32; depths, probs [B,R,P] f32;
    steps [B,R] f32; noise [B,R,S] f32.  Returns (idx, depth, dists) [B,R,S].

    ``compat`` (extension; default True) keeps the reference kernel's
    trailing-loop behaviour."""
    for n, t in (("pts_idx", pts_idx), ("min_depth", min_depth), ("max_depth", max_depth),
                 ("uniform_noise", uniform_noise), ("probs", probs), ("steps", steps)):
        _check_contig(n, t)
        _check_cuda(n, t)
    for n, t in (("min_depth", min_depth), ("max_depth", max_depth),
                 ("uniform_noise", uniform_noise), ("probs", probs), ("steps", steps)):
        _check_float(n, t)
    _check_int("pts_idx", pts_idx)
    _same_device(pts_idx, min_depth, max_depth, uniform_noise, probs, steps)
    B, R, P = pts_idx.shape
    S = uniform_noise.size(-1)
    if (min_depth.shape != pts_idx.shape or max_depth.shape != pts_idx.shape
            or probs.shape != pts_idx.shape or tuple(steps.shape) != (B, R)
            or tuple(uniform_noise.shape[:2]) != (B, R)):
        raise RuntimeError("shape mismatch: [B,R,P] bins, [B,R] steps, [B,R,S] noise")
    dev = pts_idx.device
    oi = torch.empty((B, R, S), dtype=torch.int32, device=dev)
    od = torch.empty((B, R, S), dtype=torch.float32, device=dev)
    ot = torch.empty((B, R, S), dtype=torch.float32, device=dev)
    check(lib().nsvf_inverse_cdf_sampling(_ptr(pts_idx), _ptr(min_depth), _ptr(max_depth),
                                          _ptr(uniform_noise), _ptr(probs), _ptr(steps), B, R, P,
                                          S, float(_scalar(fixed_step_size)), int(bool(compat)),
                                          _ptr(oi), _ptr(od), _ptr(ot), None, _stream(dev)),
          "inverse_cdf_sampling")
    return oi, od, ot


def build_octree(center, points, depth):
    """octree.cpp:125-136.  center [3], points [N,3] integer coordinates
    (half-voxel units; any device).  Returns (centers [T,3] i32,
    children [T,9] i32) on the device of ``center``.  The build runs on the
    host; ``center`` must hold integers (true for voxel-centre coordinates)."""
    c = center.detach().cpu()
    if c.is_floating_point():
        if not torch.equal(c, c.floor()):
            raise RuntimeError("build_octree: center must be integral")
    c = np.ascontiguousarray(c.to(torch.int64).numpy().reshape(3))
    p = np.ascontiguousarray(points.detach().cpu().to(torch.int64).numpy().reshape(-1, 3))
    L = lib()
    h = ctypes.c_void_p()
    T = ctypes.c_int()
    check(L.nsvf_build_octree(c.ctypes.data_as(ctypes.c_void_p), p.ctypes.data_as(ctypes.c_void_p),
                              p.shape[0], int(_scalar(depth)), ctypes.byref(h), ctypes.byref(T)),
          "build_octree")
    centers = np.empty((T.value, 3), np.int32)
    children = np.empty((T.value, 9), np.int32)
    check(L.nsvf_octree_take(h, centers.ctypes.data_as(ctypes.c_void_p),
                             children.ctypes.data_as(ctypes.c_void_p)), "octree_take")
    dev = center.device
    return torch.from_numpy(centers).to(dev), torch.from_numpy(children).to(dev)


def ball_intersect(*args, **kwargs):
    """Not on the sparse-voxel path (fairnr/clib/src/intersect.cpp:15-41)."""
    raise NotImplementedError("ball_intersect is not part of the sparse-voxel path")


def triangle_intersect(*args, **kwargs):
    """Not on the sparse-voxel path (fairnr/clib/src/intersect.cpp:120-146)."""
    raise NotImplementedError("triangle_intersect is not part of the sparse-voxel path")
