"""Edits of the voxel set during training: pruning and splitting
(fairnr/models/nsvf.py:143-155 -> fairnr/modules/encoder.py:605-676).

  pruning   max of relu(sigma) over a 16^3 lattice per voxel, evaluated by
            the field MLP on the matrix cores (one fused kernel per chunk of
            voxels: lattice generation, gather, MLP, per-voxel max); a voxel
            is kept when 1 - exp(-max) > th (encoder.py:610, th = 0.5).
  splitting 8 children per voxel, rebuilt corner keys and interpolated
            embeddings (geometry.py:250-274), on the device.

Both work on a DeviceScene (the resident scene) and return the edited
VoxelScene; the caller builds a new DeviceScene from it, which rebuilds the
octree, as the reference's clean_caches() does lazily.
"""
from __future__ import annotations

import ctypes
from dataclasses import replace

import numpy as np
import torch

from ._lib import check, lib
from .render import DeviceScene
from .scene import VoxelScene

PRECISIONS = {"bf16": 0, "fp32": 1, "bf16x3": 2}


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def prune_scores(sc: DeviceScene, precision="fp32", bits=16, chunk=65536, x3_chunk=None):
    """[N] float32: max over each voxel's bits^3 lattice of relu(sigma)
    (encoder.py:620-646).  FP32 MFMA by default, as the reference evaluates
    the field in FP32; "bf16" is faster and may flip voxels whose max is
    within about 1% of the threshold; "bf16x3" (split BF16, the renderer's
    served precision) keeps FP32-level accuracy at a fraction of the FP32
    time.  chunk: voxels per launch of the FP32 / BF16 kernels; x3_chunk:
    voxels per chunk of the split-BF16 path (default 2048)."""
    if precision == "bf16x3":
        return _prune_scores_x3(sc, bits, **({} if x3_chunk is None else {"chunk_voxels": x3_chunk}))
    prec = PRECISIONS[precision]
    N = sc.centers.shape[0]
    out = torch.empty(N, dtype=torch.float32, device=sc.device)
    w = sc.mlp["bf16"] if prec == 0 else sc.mlp["f32"]
    st = _stream(sc.device)
    for v0 in range(0, N, chunk):
        n = min(chunk, N - v0)
        check(lib().nsvf_field_prune(prec, sc.D, sc.H, sc.L, _p(sc.centers), _p(sc.feats),
                                     _p(sc.values), N, sc.voxel_size, v0, n, bits, _p(w),
                                     _p(sc.mlp["bias"]), _p(out[v0:]), st), "field_prune")
    return out


def _prune_scores_x3(sc: DeviceScene, bits, chunk_voxels=2048):
    """prune_scores on the split-BF16 field kernels: each chunk of voxels'
    lattice points (x = c + fp32((2i + 1 - bits) / (bits - 1)) * (vs / 2)
    per axis, i slowest, as the FP32 path places them) become a stream of
    samples -- voxel id, the point as a ray start with direction 0 and
    depth 0, so the gather's x = o + d * z is the point itself -- through
    the stream's gather and MLP kernels, then the per-voxel max of relu
    sigma.  Every point is evaluated on its own, so the chunking changes no
    value; a chunk of 2048 voxels at bits = 16 is 8.4 M samples (1.1 GB of
    operand tiles), enough that the per-chunk set-up costs little beside
    the kernels, and the direction, record and workspace buffers are built
    once for all chunks."""
    L = lib()
    dev = sc.device
    if not L.nsvf_field_bf16x3_supported(sc.D, sc.H, sc.L):
        raise RuntimeError(f"bf16x3 field not built for D={sc.D} H={sc.H} L={sc.L}")
    if int(chunk_voxels) < 1:
        raise ValueError(f"chunk_voxels {chunk_voxels}: at least 1")
    N = sc.centers.shape[0]
    out = torch.empty(N, dtype=torch.float32, device=dev)
    if N == 0:
        return out
    per = bits ** 3
    i = torch.arange(bits, dtype=torch.float32, device=dev)
    off = (2 * i + 1 - bits) / float(bits - 1)          # float32 arithmetic
    half = torch.tensor(np.float32(sc.voxel_size) * np.float32(0.5), device=dev)
    ox, oy, oz = torch.meshgrid(off, off, off, indexing="ij")
    offs = torch.stack([ox.reshape(-1), oy.reshape(-1), oz.reshape(-1)], 1) * half
    cv = min(int(chunk_voxels), N)
    mmax = cv * per
    dirs = torch.zeros((mmax, 3), dtype=torch.float32, device=dev)
    rec = torch.zeros((mmax, 4), dtype=torch.int32, device=dev)
    local = torch.arange(cv, dtype=torch.int32, device=dev).repeat_interleave(per)
    rec[:, 3] = torch.arange(mmax, dtype=torch.int32, device=dev)
    pts = torch.empty((cv, per, 3), dtype=torch.float32, device=dev)
    field = torch.empty((mmax, 4), dtype=torch.float32, device=dev)
    ws_bytes = int(L.nsvf_field_workspace_bytes(2, sc.D, sc.H, sc.L, mmax))
    ws = torch.empty((ws_bytes + 15) // 16 * 4, dtype=torch.int32, device=dev)
    m_dev = torch.full((1,), mmax, dtype=torch.int32, device=dev)
    st = _stream(dev)
    w = sc.mlp["bf16x3"]
    for v0 in range(0, N, cv):
        n = min(cv, N - v0)
        m = n * per
        if n < cv:
            m_dev.fill_(m)
        torch.add(sc.centers[v0:v0 + n, None, :], offs[None], out=pts[:n])
        torch.add(local[:m], v0, out=rec[:m, 0])
        check(L.nsvf_field_embed_bf16x3(sc.D, sc.H, sc.L, _p(sc.centers), _p(sc.feats),
                                        _p(sc.values), sc.voxel_size, _p(pts), _p(dirs), _p(rec),
                                        None, _p(m_dev), m, _p(ws), ws_bytes, st), "field_embed")
        check(L.nsvf_field_mlp_bf16x3(sc.D, sc.H, sc.L, _p(ws), ws_bytes, _p(m_dev), m, _p(w),
                                      _p(sc.mlp["bias"]), _p(field), None, st), "field_mlp")
        torch.amax(field[:m, 0].view(n, per).clamp(min=0), 1, out=out[v0:v0 + n])
    return out


def pruning(sc: DeviceScene, th=0.5, precision="fp32", bits=16, train_stats=False,
            max_voxel_probs=None, delta=None):
    """keep [N] bool (encoder.py:605-615).  From the field: 1 - min(exp(-relu
    sigma)) > th over each voxel's lattice, computed from the max as
    1 - exp(-max) > th (exp is monotone).  delta: the opacity over an
    interval of that length instead, 1 - exp(-max * delta) > th (the rule
    "prune where max(1 - exp(-sigma delta)) < th"; the reference's own rule
    has no interval, i.e. delta = 1).  With train_stats: max_voxel_probs > th,
    the statistics a Renderer gathers in its ``track`` tensor while rendering
    the training views (encoder.py:594-603)."""
    if train_stats:
        if max_voxel_probs is None or max_voxel_probs.numel() != sc.scene.num_voxels:
            raise ValueError("train_stats pruning needs max_voxel_probs [num_voxels]")
        return max_voxel_probs > th
    vmax = prune_scores(sc, precision, bits)
    if delta is not None:
        if not float(delta) > 0.0:
            raise ValueError(f"delta {delta}: positive")
        vmax = vmax * float(delta)
    return (1 - torch.exp(-vmax)) > th


def prune_scene(scene: VoxelScene, keep) -> VoxelScene:
    """The scene with only the kept voxels (encoder.py:617 masks them out of
    points and feats; the key table and embeddings stay as they are)."""
    keep = np.asarray(keep.cpu().numpy() if torch.is_tensor(keep) else keep, bool)
    return replace(scene, points=np.ascontiguousarray(scene.points[keep]),
                   feats=np.ascontiguousarray(scene.feats[keep]), meta=dict(scene.meta))


def split_scene(scene: VoxelScene, device) -> VoxelScene:
    """encoder.py:656-676 + nsvf.py:148-155 on the device: 8 children per
    voxel, new corner keys and embeddings, voxel size halved, max hits x1.5
    (kept as a float, as the reference's buffer is; the intersection uses its
    integer part).

    As in the reference, the split starts from the encoder's precomputed
    centres, which carry the gather's +vs/10 shift in x (encoder.py:383,
    659-661), so the children's centres keep that shift."""
    dev = torch.device(device)
    L = lib()
    vs = np.float32(scene.voxel_size)
    pts = np.array(scene.points, np.float32, copy=True)
    pts[:, 0] += vs / np.float32(10)
    N = pts.shape[0]
    D = scene.values.shape[1]
    st = _stream(dev)
    points = torch.from_numpy(pts).to(dev)
    feats = torch.from_numpy(np.ascontiguousarray(scene.feats, np.int32)).to(dev)
    values = torch.from_numpy(np.ascontiguousarray(scene.values, np.float32)).to(dev)
    nbytes = int(L.nsvf_split_workspace_bytes(N, st))
    if nbytes < 0:
        check(-1, "split_workspace_bytes")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    new_points = torch.empty((8 * N, 3), dtype=torch.float32, device=dev)
    new_feats = torch.empty((8 * N, 8), dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    check(L.nsvf_split_voxels(_p(points), N, float(vs), _p(ws), nbytes, _p(new_points),
                              _p(new_feats), _p(count), st), "split_voxels")
    K = int(count.item())
    new_keys = torch.empty((K, 3), dtype=torch.int64, device=dev)
    new_values = torch.empty((K, D), dtype=torch.float32, device=dev)
    check(L.nsvf_split_values(_p(points), _p(feats), _p(values), N, D, float(vs), _p(ws), nbytes,
                              K, _p(new_keys), _p(new_values), st), "split_values")
    half = float(vs * np.float32(0.5))
    return replace(scene, points=new_points.cpu().numpy(), voxel_size=half,
                   max_hits=scene.max_hits * 1.5, keys=new_keys.cpu().numpy(),
                   feats=new_feats.cpu().numpy().astype(np.int64),
                   values=new_values.cpu().numpy(), meta=dict(scene.meta))
