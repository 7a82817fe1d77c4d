"""CPU oracle for the NSVF sparse-voxel rendering path.

TEST INFRASTRUCTURE ONLY.  Tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module to check (or time) the
HIP library; no product code path imports it.

Two layers:

* FP32 restatements, bit-for-bit in intent with the reference:
  - the four native ops and the octree builder, in C++ (``cpu_ref.cpp``,
    loaded here through ctypes);
  - the Python glue around them, in NumPy float32 with a fixed evaluation
    order (left-to-right sums, stable sorts).
* FP64 restatements of the continuous parts (trilinear gather, field MLP,
  compositing) used to state the error of every floating-point output.

Reference files restated (line numbers in each docstring):
  fairnr/clib/__init__.py, fairnr/clib/src/*.cu|*.cpp, fairnr/data/geometry.py,
  fairnr/modules/encoder.py, fairnr/modules/renderer.py, fairnr/models/nsvf.py,
  fairnr/models/nerf.py, fairnr/models/fairnr_model.py.

Choices where the reference leaves the result to a library's internal order:
  * ``sort`` of hit depths is stable (ties keep slot order);
  * ``sum`` of segment lengths and of compositing terms runs left to right;
  * the corner sum of the gather runs over corners 0..7;
  * ``cumsum`` of free energy runs left to right;
  * the octree residual is the float64 mean of (p - coord*h), rounded to fp32.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

MAX_DEPTH = 10000.0          # fairnr/clib/__init__.py:36
FREE_ENERGY_SCALE = 7.0      # fairnr/modules/renderer.py:121
BG_DEPTH = 5.0               # fairnr/modules/implicit.py:19 (background_depth)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load ``liboracle_cpu.so`` (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle_cpu.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle library missing: run `make -C oracle`")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i, f = ctypes.c_int, ctypes.c_float
    L.orc_aabb_intersect.argtypes = [i, i, i, P, P, P, f, i, P, P, P]
    L.orc_aabb_intersect.restype = None
    L.orc_svo_intersect.argtypes = [i, i, i, P, P, P, P, f, i, P, P, P]
    L.orc_svo_intersect.restype = i
    L.orc_octree_build.argtypes = [P, P, i, i]
    L.orc_octree_build.restype = P
    L.orc_octree_num_nodes.argtypes = [P]
    L.orc_octree_num_nodes.restype = i
    L.orc_octree_num_leaves.argtypes = [P]
    L.orc_octree_num_leaves.restype = i
    L.orc_octree_copy.argtypes = [P, P, P]
    L.orc_octree_copy.restype = None
    L.orc_octree_free.argtypes = [P]
    L.orc_octree_free.restype = None
    L.orc_inverse_cdf_sampling.argtypes = [i, i, i, i, f, P, P, P, P, P, P, i, P, P, P, P]
    L.orc_inverse_cdf_sampling.restype = None
    L.orc_uniform_ray_sampling.argtypes = [i, i, i, i, f, P, P, P, P, P, P, P, P]
    L.orc_uniform_ray_sampling.restype = None
    L.orc_aabb_intersect_f64.argtypes = [i, i, P, P, P, ctypes.c_double, i, P, P, P]
    L.orc_aabb_intersect_f64.restype = i
    _LIB = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# Native ops, op level (the `_ext` boundary: fairnr/clib/src/binding.cpp:11-21)

def aabb_intersect(ray_start, ray_dir, points, voxelsize, n_max):
    """``_ext.aabb_intersect``: intersect.cpp:49-75, intersect_gpu.cu:125-167.

    ray_start, ray_dir: [B,M,3] f32; points: [B,N,3] f32.
    Returns idx [B,M,n_max] i32 (-1 pad), min/max depth f32 (0 pad)."""
    rs, rd, pts = _c(ray_start, np.float32), _c(ray_dir, np.float32), _c(points, np.float32)
    B, M = rs.shape[:2]
    N = pts.shape[1]
    n_max = int(n_max)
    idx = np.zeros((B, M, n_max), np.int32)
    lo = np.zeros((B, M, n_max), np.float32)
    hi = np.zeros((B, M, n_max), np.float32)
    lib().orc_aabb_intersect(B, M, N, _p(rs), _p(rd), _p(pts), float(voxelsize), n_max,
                             _p(idx), _p(lo), _p(hi))
    return idx, lo, hi


def svo_intersect(ray_start, ray_dir, points, children, voxelsize, n_max):
    """``_ext.svo_intersect``: intersect.cpp:84-112, intersect_gpu.cu:170-237.

    points: [B,T,3] f32 node centres; children: [B,T,9] i32.  Root is node T-1."""
    rs, rd = _c(ray_start, np.float32), _c(ray_dir, np.float32)
    pts, ch = _c(points, np.float32), _c(children, np.int32)
    B, M = rs.shape[:2]
    T = pts.shape[1]
    n_max = int(n_max)
    idx = np.zeros((B, M, n_max), np.int32)
    lo = np.zeros((B, M, n_max), np.float32)
    hi = np.zeros((B, M, n_max), np.float32)
    ovf = lib().orc_svo_intersect(B, M, T, _p(rs), _p(rd), _p(pts), _p(ch), float(voxelsize),
                                  n_max, _p(idx), _p(lo), _p(hi))
    if ovf:
        raise RuntimeError(f"svo_intersect: DFS stack exceeded 256 entries on {ovf} rays")
    return idx, lo, hi


def build_octree(center, points, depth):
    """``_ext.build_octree``: octree.cpp:125-136.

    center: [3] int; points: [N,3] int (half-voxel units).  Returns
    centers [T,3] i32 and children [T,9] i32 (-1 = none, [8] = 1<<(depth+1))."""
    c = _c(np.asarray(center).reshape(3), np.int64)
    pts = _c(points, np.int64).reshape(-1, 3)
    if int(depth) < 0:
        raise RuntimeError("build_octree: depth must be >= 0")
    L = lib()
    h = L.orc_octree_build(_p(c), _p(pts), pts.shape[0], int(depth))
    try:
        T = L.orc_octree_num_nodes(h)
        centers = np.zeros((T, 3), np.int32)
        children = np.zeros((T, 9), np.int32)
        L.orc_octree_copy(h, _p(centers), _p(children))
    finally:
        L.orc_octree_free(h)
    return centers, children


def inverse_cdf_sampling_op(pts_idx, min_depth, max_depth, noise, probs, steps,
                            fixed_step_size, compat=True):
    """``_ext.inverse_cdf_sampling``: sample.cpp:58-95, sample_gpu.cu:108-202.

    Inputs [B,R,P]; steps [B,R]; noise [B,R,S].  Outputs [B,R,S] with idx -1,
    depth 0 and dists 0 in unused slots.  Returns (idx, depth, dists, flags)."""
    pi = _c(pts_idx, np.int32)
    mn, mx = _c(min_depth, np.float32), _c(max_depth, np.float32)
    nz, pr, st = _c(noise, np.float32), _c(probs, np.float32), _c(steps, np.float32)
    B, R, P = pi.shape
    S = nz.shape[-1]
    oi = np.full((B, R, S), -1, np.int32)
    od = np.zeros((B, R, S), np.float32)
    ot = np.zeros((B, R, S), np.float32)
    flags = np.zeros(4, np.int64)
    lib().orc_inverse_cdf_sampling(B, R, P, S, float(fixed_step_size), _p(pi), _p(mn), _p(mx),
                                   _p(nz), _p(pr), _p(st), int(bool(compat)), _p(oi), _p(od),
                                   _p(ot), _p(flags))
    return oi, od, ot, flags


def uniform_ray_sampling_op(pts_idx, min_depth, max_depth, noise, step_size, max_steps):
    """``_ext.uniform_ray_sampling``: sample.cpp:23-55, sample_gpu.cu:15-106.

    Returns (idx, depth, dists, flags); outputs [B,R,max_steps]."""
    pi = _c(pts_idx, np.int32)
    mn, mx = _c(min_depth, np.float32), _c(max_depth, np.float32)
    nz = _c(noise, np.float32)
    B, R, P = pi.shape
    S = int(max_steps)
    assert nz.shape == (B, R, S)
    oi = np.full((B, R, S), -1, np.int32)
    od = np.zeros((B, R, S), np.float32)
    ot = np.zeros((B, R, S), np.float32)
    flags = np.zeros(4, np.int64)
    lib().orc_uniform_ray_sampling(B, R, P, S, float(step_size), _p(pi), _p(mn), _p(mx), _p(nz),
                                   _p(oi), _p(od), _p(ot), _p(flags))
    return oi, od, ot, flags


# ---------------------------------------------------------------------------
# Python-level wrappers (fairnr/clib/__init__.py)

def aabb_ray_intersect(voxelsize, n_max, points, ray_start, ray_dir):
    """``aabb_ray_intersect``: clib/__init__.py:58-95.  The wrapper's regrouping
    (G blocks of K rays, repeated pad rays, trimmed afterwards) does not change
    any ray's result, so the op runs once over [S, N] rays."""
    return aabb_intersect(ray_start, ray_dir, points, voxelsize, n_max)


def svo_ray_intersect(voxelsize, n_max, points, children, ray_start, ray_dir):
    """``svo_ray_intersect``: clib/__init__.py:98-135 (same note as above)."""
    return svo_intersect(ray_start, ray_dir, points, children, voxelsize, n_max)


def inverse_cdf_sampling(pts_idx, min_depth, max_depth, probs, steps, fixed_step_size=-1.0,
                         deterministic=True, noise=None, compat=True):
    """``inverse_cdf_sampling``: clib/__init__.py:231-300.

    Rows are padded to a multiple of G=200 with copies of row 0, regrouped to
    [G, H/G, P] and processed 800 columns (4*G) per op call; the results are
    concatenated, trimmed back to N rows and to the longest row.  In compat
    mode the grouping decides which ray is "ray 0 of the row" and "the next
    ray" for the reference's trailing-loop quirks.

    noise: [N, max_steps] (only when not deterministic).  Returns
    (idx [N,L] i32, depth [N,L] f32, dists [N,L] f32, max_steps, flags)."""
    pts_idx = _c(pts_idx, np.int32)
    N, P = pts_idx.shape
    G = 200
    H = int(math.ceil(N / G)) * G
    arrs = [pts_idx, _c(min_depth, np.float32), _c(max_depth, np.float32), _c(probs, np.float32)]
    st = _c(steps, np.float32)
    if H > N:
        arrs = [np.concatenate([a, np.repeat(a[:1], H - N, 0)], 0) for a in arrs]
        st = np.concatenate([st, np.repeat(st[:1], H - N, 0)], 0)
    Kc = H // G
    arrs = [a.reshape(G, Kc, P) for a in arrs]
    st = st.reshape(G, Kc)
    max_steps = int(np.ceil(st).astype(np.int64).max()) + P
    if deterministic:
        nz = np.full((G, Kc, max_steps), 0.5, np.float32)
    else:
        assert noise is not None and noise.shape[1] >= max_steps
        nz_full = _c(noise[:, :max_steps], np.float32)
        if H > N:
            nz_full = np.concatenate([nz_full, np.repeat(nz_full[:1], H - N, 0)], 0)
        nz = nz_full.reshape(G, Kc, max_steps)
    chunk = 4 * G
    outs = []
    flags = np.zeros(4, np.int64)
    for i in range(0, Kc, chunk):
        sl = slice(i, i + chunk)
        oi, od, ot, fl = inverse_cdf_sampling_op(arrs[0][:, sl], arrs[1][:, sl], arrs[2][:, sl],
                                                 nz[:, sl], arrs[3][:, sl], st[:, sl],
                                                 fixed_step_size, compat)
        outs.append((oi, od, ot))
        flags += fl
    si = np.concatenate([o[0] for o in outs], 1).reshape(H, -1)[:N]
    sd = np.concatenate([o[1] for o in outs], 1).reshape(H, -1)[:N]
    sx = np.concatenate([o[2] for o in outs], 1).reshape(H, -1)[:N]
    max_len = int((si != -1).sum(-1).max()) if N > 0 else 0
    return si[:, :max_len], sd[:, :max_len], sx[:, :max_len], max_steps, flags


def pcg_hash(v):
    """PCG hash (Jarzynski and Olano, "Hash functions for GPU rendering",
    JCGT 9(3) 2020) of uint32 values, modulo 2^32."""
    v = np.asarray(v, np.uint32)
    with np.errstate(over="ignore"):
        state = v * np.uint32(747796405) + np.uint32(2891336453)
        word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def hash_noise(seed, rows, n_steps):
    """The samplers' training-mode noise [len(rows), n_steps] float32: step k
    of frame ray r draws u = (pcg(k + pcg(r + pcg(seed))) >> 8) / 2^24,
    clamped to [0.001, 0.999] as clib/__init__.py:256 clamps its uniform_()
    draws (the reference's torch generator cannot be reproduced; this
    counter-based stream is the library's own, for any seed in [0, 2^32))."""
    with np.errstate(over="ignore"):
        key = pcg_hash(np.asarray(rows, np.int64).astype(np.uint32) + pcg_hash(np.uint32(seed)))
        k = np.arange(n_steps, dtype=np.uint32)
        h = pcg_hash(k[None, :] + key[:, None])
    u = (h >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)
    return np.minimum(np.maximum(u, np.float32(0.001)), np.float32(0.999))


def pass_seed(seed, k):
    """Seed of sampling pass k (0 coarse, 1 fine) of a view rendered with
    noise seed `seed`: seed + k * 0x9E3779B9 modulo 2^32."""
    return (int(seed) + k * 0x9E3779B9) & 0xFFFFFFFF


def uniform_ray_sampling(pts_idx, min_depth, max_depth, step_size, max_ray_length,
                         deterministic=True, noise=None):
    """``uniform_ray_sampling``: clib/__init__.py:178-228.

    max_steps = int(max_ray_length / step_size) + 2P.  The G=256 regrouping is
    contiguous, so every ray sees the same neighbours as in one [1,N,P] call
    and the op runs once.  Returns (idx, depth, dists, max_steps, flags)."""
    pts_idx = _c(pts_idx, np.int32)
    N, P = pts_idx.shape
    max_steps = int(max_ray_length / step_size) + P * 2
    if deterministic:
        nz = np.full((1, N, max_steps), 0.5, np.float32)
    else:
        nz = _c(noise[:, :max_steps], np.float32)[None]
    oi, od, ot, fl = uniform_ray_sampling_op(pts_idx[None], min_depth[None], max_depth[None], nz,
                                             step_size, max_steps)
    oi, od, ot = oi[0], od[0], ot[0]
    max_len = int((oi != -1).sum(-1).max()) if N > 0 else 0
    return oi[:, :max_len], od[:, :max_len], ot[:, :max_len], max_steps, fl


# ---------------------------------------------------------------------------
# Geometry helpers (fairnr/data/geometry.py)

def offset_points(point_xyz, quarter_voxel=1.0, offset_only=False, bits=2):
    """geometry.py:229-238: lattice offsets (2i+1-bits)/(bits-1) * quarter_voxel,
    i-major (x slowest, z fastest: meshgrid 'ij').  Integer inputs give float32
    outputs with integer offsets, as the reference's type_as(long) does."""
    c = np.arange(1, 2 * bits, 2)
    ox, oy, oz = np.meshgrid(c, c, c, indexing="ij")
    off_i = np.stack([ox.reshape(-1), oy.reshape(-1), oz.reshape(-1)], 1) - bits
    p = np.asarray(point_xyz)
    if p.dtype.kind in "iu":
        off = (off_i / float(bits - 1)).astype(np.int64).astype(np.float32) * np.float32(quarter_voxel)
        if offset_only:
            return off
        return p[:, None, :].astype(np.float32) + off[None]
    dt = p.dtype.type
    off = (off_i.astype(p.dtype) / dt(bits - 1)) * dt(quarter_voxel)
    if offset_only:
        return off
    return p[:, None, :] + off[None]


def discretize_points(voxel_points, voxel_size):
    """geometry.py:241-247: integer coordinates relative to the minimum point,
    and the mean residual (float64 mean, rounded to fp32)."""
    vp = np.asarray(voxel_points, np.float32)
    vs = np.float32(voxel_size)
    mn = vp.min(0, keepdims=True)
    coords = np.round((vp - mn) / vs).astype(np.int64)
    residual = (vp.astype(np.float64) - (coords.astype(np.float32) * vs).astype(np.float64)).mean(0, keepdims=True)
    return coords, residual.astype(np.float32)


def build_easy_octree(points, half_voxel):
    """geometry.py:320-327: octree over voxel centres in half-voxel units.

    depth = ceil(log2(max range)) - 1 (0 for a single voxel, where the
    reference is undefined); centre = (max + min) / 2."""
    coords, residual = discretize_points(points, half_voxel)
    ranges = coords.max(0) - coords.min(0)
    rmax = int(ranges.max())
    depth = int(math.ceil(math.log2(rmax))) - 1 if rmax > 0 else 0
    depth = max(depth, 0)
    center = (coords.max(0) + coords.min(0)) // 2
    centers, children = build_octree(center, coords, depth)
    centers_f = centers.astype(np.float32) * np.float32(half_voxel) + residual
    return centers_f.astype(np.float32), children


def voxel_keys(points, voxel_size):
    """encoder.py:269-276: corner keys of every voxel, unique and sorted, and
    each voxel's 8 corner ids (corner order = offset_points order)."""
    coords, _ = discretize_points(points, np.float32(voxel_size) * np.float32(0.5))
    keys0 = offset_points(coords, 1.0).reshape(-1, 3).astype(np.int64)
    keys, inv = np.unique(keys0, axis=0, return_inverse=True)
    return keys, inv.reshape(-1, 8).astype(np.int64)


def shift_centers(points, voxel_size):
    """encoder.py:383: x += voxel_size / 10 (fp32)."""
    p = np.array(points, np.float32, copy=True)
    p[:, 0] += np.float32(voxel_size) / np.float32(10)
    return p


# ---------------------------------------------------------------------------
# Post-intersection glue (encoder.py:518-535, nsvf.py:65-74)

def sort_hits(idx, min_depth, max_depth):
    """encoder.py:519-524: pads -> MAX_DEPTH, stable sort by entry depth."""
    idx = np.asarray(idx, np.int32)
    pad = idx == -1
    mn = np.where(pad, np.float32(MAX_DEPTH), min_depth).astype(np.float32)
    mx = np.where(pad, np.float32(MAX_DEPTH), max_depth).astype(np.float32)
    order = np.argsort(mn, axis=-1, kind="stable")
    mn = np.take_along_axis(mn, order, -1)
    mx = np.take_along_axis(mx, order, -1)
    idx = np.take_along_axis(idx, order, -1)
    hits = (idx != -1).any(-1)
    return idx, mn, mx, hits


def probs_steps(idx, min_depth, max_depth, step_size):
    """nsvf.py:68-74: probs = seg / sum(seg), steps = sum(seg) / step_size, in
    fp32 with a left-to-right sum."""
    d = np.where(idx == -1, np.float32(0), (max_depth - min_depth).astype(np.float32)).astype(np.float32)
    tot = np.zeros(d.shape[:-1], np.float32)
    for k in range(d.shape[-1]):
        tot = (tot + d[..., k]).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        probs = (d / tot[..., None]).astype(np.float32)
        steps = (tot / np.float32(step_size)).astype(np.float32)
    return probs, steps


def sample_tail(idx, depth, dists):
    """encoder.py:547-549: clamp dists at 0; pads -> MAX_DEPTH depth, 0 dist."""
    dists = np.maximum(dists, np.float32(0)).astype(np.float32)
    pad = idx == -1
    depth = np.where(pad, np.float32(MAX_DEPTH), depth).astype(np.float32)
    dists = np.where(pad, np.float32(0), dists).astype(np.float32)
    return depth, dists


# ---------------------------------------------------------------------------
# Trilinear gather (encoder.py:582-590, geometry.py:195-200)

def trilinear_gather(sample_idx, xyz, centers, feats, values, voxel_size, dtype=np.float64):
    """emb = sum_c w_c * values[feats[idx][c]], p = (x - centre)/vs + 0.5,
    w_c = prod_axis (p if bit else 1-p), corner c = 4x + 2y + z.

    Evaluated in ``dtype`` (float64 for the oracle; float32 restates the
    reference's FP32 arithmetic with the corner sum in corner order)."""
    dt = np.dtype(dtype).type
    idx = np.asarray(sample_idx, np.int64)
    x = np.asarray(xyz, dtype)
    c = np.asarray(centers, dtype)[idx]
    p = ((x - c) / dt(voxel_size) + dt(0.5)).astype(dtype)
    E = np.asarray(values, dtype)[np.asarray(feats)[idx]]          # [M,8,D]
    out = np.zeros((x.shape[0], E.shape[-1]), dtype)
    for corner in range(8):
        bx, by, bz = (corner >> 2) & 1, (corner >> 1) & 1, corner & 1
        wx = p[:, 0] if bx else (dt(1) - p[:, 0])
        wy = p[:, 1] if by else (dt(1) - p[:, 1])
        wz = p[:, 2] if bz else (dt(1) - p[:, 2])
        w = ((wx * wy).astype(dtype) * wz).astype(dtype)
        out = (out + (w[:, None] * E[:, corner]).astype(dtype)).astype(dtype)
    return out


def trilinear_weights(sample_idx, xyz, centers, voxel_size, dtype=np.float64):
    """Per-sample corner weights [M,8] (same formula as trilinear_gather)."""
    dt = np.dtype(dtype).type
    idx = np.asarray(sample_idx, np.int64)
    x = np.asarray(xyz, dtype)
    c = np.asarray(centers, dtype)[idx]
    p = ((x - c) / dt(voxel_size) + dt(0.5)).astype(dtype)
    W = np.zeros((x.shape[0], 8), dtype)
    for corner in range(8):
        bx, by, bz = (corner >> 2) & 1, (corner >> 1) & 1, corner & 1
        wx = p[:, 0] if bx else (dt(1) - p[:, 0])
        wy = p[:, 1] if by else (dt(1) - p[:, 1])
        wz = p[:, 2] if bz else (dt(1) - p[:, 2])
        W[:, corner] = (wx * wy) * wz
    return W


# ---------------------------------------------------------------------------
# Field MLP (the configs' own MLPs: L hidden ReLU layers, then a 4-wide head
# giving sigma and RGB logits; RGB through a sigmoid as with min_color=0 in
# fairnr/modules/field.py:275-276)

@dataclass
class MLPWeights:
    weights: list   # [out, in] matrices, hidden layers then head
    biases: list


def mlp_forward(x, w: MLPWeights, dtype=np.float64):
    h = np.asarray(x, dtype)
    n = len(w.weights)
    for i, (W, b) in enumerate(zip(w.weights, w.biases)):
        h = h @ np.asarray(W, dtype).T + np.asarray(b, dtype)
        if i < n - 1:
            h = np.maximum(h, 0)
    sigma = h[:, 0]
    rgb = 1.0 / (1.0 + np.exp(-h[:, 1:4]))
    return sigma.astype(dtype), rgb.astype(dtype)


# ---------------------------------------------------------------------------
# Compositing (renderer.py:117-124, :193-218) and background (nsvf.py:94-101)

def composite(sigma, rgb, depth, dists, idx, dtype=np.float64):
    """Front-to-back compositing of padded [R,S] samples.

    fe = relu(sigma) * dist * 7 (0 at pads); T_k = exp(-sum_{j<k} fe_j);
    p_k = (1 - exp(-fe_k)) T_k; depth = sum p z; missed = 1 - sum p;
    rgb = sum p c.  Returns (rgb [R,3], depth [R], missed [R], probs [R,S])."""
    sigma = np.asarray(sigma, dtype)
    valid = np.asarray(idx) != -1
    fe = np.where(valid, np.maximum(sigma, 0) * np.asarray(dists, dtype) * FREE_ENERGY_SCALE, 0).astype(dtype)
    shifted = np.concatenate([np.zeros_like(fe[:, :1]), fe[:, :-1]], 1)
    a = 1 - np.exp(-fe)
    T = np.exp(-np.cumsum(shifted, 1))
    probs = a * T
    z = np.asarray(depth, dtype)
    out_depth = (z * probs).sum(-1)
    missed = 1 - probs.sum(-1)
    col = (np.where(valid[..., None], np.asarray(rgb, dtype), 0) * probs[..., None]).sum(1)
    return col, out_depth, missed, probs


def march_chunks(idx, sigma, dists, tolerance, chunk_size):
    """Which samples the reference's chunked ray marching evaluates
    (renderer.py:136-174, with forward():234-247 splitting the rows).

    Rows (hit rays) go in groups of chunk_size; in a group, the columns of
    the padded [N,S] samples go in chunks: a chunk ends before the column
    whose count of hits (idx != -1, rows not yet stopped) would take the
    chunk's total past chunk_size (a chunk holds at least one column).
    After each chunk that evaluated something, every row adds the float32
    sum (left to right) of its chunk's free energies relu(sigma) * dist * 7
    (0 where not evaluated) to its running total, and rows whose total
    exceeds float32(-log(tolerance)) stop: their later samples are not
    evaluated (their field is 0) and no longer count toward chunk sizes.

    sigma: [N,S] field sigma (cast to float32 for the decisions).  Returns
    (evaluated [N,S] bool, stopped [N] bool, margin [N] float64: the least
    |total - threshold| / threshold over the row's chunk ends, to tell rows
    whose decision another implementation's rounding could flip)."""
    idx = np.asarray(idx)
    N, S = idx.shape
    sig = np.asarray(sigma, np.float64).astype(np.float32)
    dist = np.asarray(dists, np.float32)
    fe = ((np.maximum(sig, np.float32(0)) * dist).astype(np.float32) *
          np.float32(FREE_ENERGY_SCALE)).astype(np.float32)
    tol = np.float32(-math.log(tolerance)) if tolerance > 0 else None
    evaluated = np.zeros((N, S), bool)
    stopped = np.zeros(N, bool)
    margin = np.full(N, np.inf)
    for r0 in range(0, N, chunk_size):
        rows = slice(r0, min(N, r0 + chunk_size))
        hits = (idx[rows] != -1).astype(np.int64)
        acc = np.zeros(hits.shape[0], np.float32)
        early = np.zeros(hits.shape[0], bool)
        size, start = 0, 0
        for i in range(S + 1):
            if (i == S or size + hits[:, i].sum() > chunk_size) and i > start:
                mask = (idx[rows, start:i] != -1) & ~early[:, None]
                if mask.any():
                    evaluated[rows, start:i] = mask
                    part = np.zeros(hits.shape[0], np.float32)
                    for j in range(i - start):
                        part = (part + np.where(mask[:, j], fe[rows, start + j], 0)).astype(np.float32)
                    acc = (acc + part).astype(np.float32)
                    if tol is not None:
                        live = ~early
                        m = np.abs(acc.astype(np.float64) - float(tol)) / float(tol)
                        margin[rows][live] = np.minimum(margin[rows][live], m[live])
                        early = acc > tol
                        hits[early] = 0
                start, size = i, 0
            if i < S:
                size += hits[:, i].sum()
        stopped[rows] = early
    return evaluated, stopped, margin


def composite_backward(sigma, rgb, depth, dists, idx, g_rgb, g_depth, g_missed, dtype=np.float64):
    """Gradients of composite()'s (rgb, depth, missed) with respect to sigma
    [R,S] and rgb [R,S,3], as autograd forms them through renderer.py:
    119-121,195-218.  With g_p(k) = <g_rgb, c_k> + g_depth z_k - g_missed:
      d/dfe_k = g_p(k) T_k exp(-fe_k) - sum_{j>k} g_p(j) p_j
      d/dsigma_k = d/dfe_k * 7 dist_k where sigma_k > 0 (relu), else 0
      d/dc_k = p_k g_rgb.
    Returns (g_sigma [R,S], g_rgb_samples [R,S,3]); zero at pads."""
    sigma = np.asarray(sigma, dtype)
    dists = np.asarray(dists, dtype)
    z = np.asarray(depth, dtype)
    c = np.asarray(rgb, dtype)
    valid = np.asarray(idx) != -1
    fe = np.where(valid, np.maximum(sigma, 0) * dists * FREE_ENERGY_SCALE, 0).astype(dtype)
    excl = np.cumsum(np.concatenate([np.zeros_like(fe[:, :1]), fe[:, :-1]], 1), 1)
    T = np.exp(-excl)
    e = np.exp(-fe)
    p = (1 - e) * T
    gr = np.asarray(g_rgb, dtype)
    gp = np.where(valid, (c * gr[:, None, :]).sum(-1) + z * np.asarray(g_depth, dtype)[:, None], 0) \
        - np.asarray(g_missed, dtype)[:, None]
    u = gp * p
    later = np.cumsum(u[:, ::-1], 1)[:, ::-1] - u          # sum over j > k
    dfe = gp * T * e - later
    g_sigma = np.where(valid & (sigma > 0), dfe * dists * FREE_ENERGY_SCALE, 0)
    g_c = np.where(valid[..., None], p[..., None] * gr[:, None, :], 0)
    return g_sigma, g_c


def trilinear_backward(sample_idx, xyz, centers, feats, num_keys, grad_out, voxel_size,
                       dtype=np.float64):
    """Gradient of trilinear_gather with respect to the embedding table: the
    adjoint scatter-add grad[feats[v][c]] += w_c * grad_out[m] that torch's
    F.embedding backward performs for encoder.py:588-590.  Samples with id -1
    contribute nothing.  Returns [num_keys, D]."""
    idx = np.asarray(sample_idx, np.int64)
    g = np.asarray(grad_out, dtype)
    D = g.shape[1]
    out = np.zeros((num_keys, D), dtype)
    ok = idx >= 0
    if not ok.any():
        return out
    W = trilinear_weights(idx[ok], np.asarray(xyz)[ok], centers, voxel_size, dtype=dtype)
    keys = np.asarray(feats, np.int64)[idx[ok]]                   # [M,8]
    np.add.at(out, keys.reshape(-1), (W[:, :, None] * g[ok][:, None, :]).reshape(-1, D))
    return out


def track_voxel_probs(idx, probs, num_voxels, prev=None):
    """encoder.py:594-603: per ray, the sum of the sample probabilities in
    each voxel (scatter_add into voxel columns; -1 samples go to a dropped
    column), then the max over rays, then the max with the running values.
    Sums run left to right in sample order, in float32."""
    out = np.zeros(num_voxels, np.float32) if prev is None else np.array(prev, np.float32)
    idx = np.asarray(idx)
    probs = np.asarray(probs, np.float32)
    for r in range(idx.shape[0]):
        sums = {}
        for s in range(idx.shape[1]):
            v = int(idx[r, s])
            if v < 0:
                continue
            sums[v] = np.float32(sums.get(v, np.float32(0)) + probs[r, s])
        for v, x in sums.items():
            out[v] = max(out[v], x)
    return out


def camera_uv(H, W, h, w):
    """data_utils.py:57-65 ``get_uv``: [2, h, w] float32 pixel coordinates,
    u = column * float32(W / w), v = row * float32(H / h)."""
    uv = np.flip(np.mgrid[0:h, 0:w], axis=0).astype(np.float32)
    uv[0] = uv[0] * np.float32(W / w)
    uv[1] = uv[1] * np.float32(H / h)
    return uv


def camera_rays(H, W, h, w, fx, fy, cx, cy, inv_RT, dtype=np.float64):
    """Rays of one view (renderer.py:91-112, reader.py:92-119,
    geometry.py:76-89,116-122): uv from ``camera_uv``; x = (u - cx) / fx,
    y = (v - cy) / fy, z = 1 (``uv2cam``); rt = inv_RT [x, y, 1, 1]
    (``cam2world``); d = rt - t; dir = d / (|d| + 1e-8) (``normalize``).
    Returns ray_start, ray_dir [h*w, 3] in row-major pixel order."""
    uv = camera_uv(H, W, h, w).reshape(2, -1).astype(dtype)
    M = np.asarray(inv_RT, dtype)[:3, :4]
    x = (uv[0] - dtype(cx)) / dtype(fx)
    y = (uv[1] - dtype(cy)) / dtype(fy)
    cam = np.stack([x, y, np.ones_like(x), np.ones_like(x)])
    rt = M @ cam
    t = M[:, 3:4]
    d = rt - t
    n = np.sqrt((d * d).sum(0, keepdims=True)) + 1e-8
    start = np.broadcast_to(t.T, (uv.shape[1], 3))
    return np.ascontiguousarray(start), np.ascontiguousarray((d / n).T)


def fill_background(hits, rgb, depth, missed, bg=(1.0, 1.0, 1.0), bg_depth=BG_DEPTH, dtype=np.float64):
    """nsvf.py:89-101 with geometry.py:303-317: scatter hit rays back to the
    full frame (missed=1, rgb=0, depth=0 elsewhere), then add the background."""
    R = hits.shape[0]
    full_rgb = np.zeros((R, 3), dtype)
    full_depth = np.zeros(R, dtype)
    full_missed = np.ones(R, dtype)
    full_rgb[hits] = rgb
    full_depth[hits] = depth
    full_missed[hits] = missed
    full_rgb = full_rgb + full_missed[:, None] * np.asarray(bg, dtype)[None]
    full_depth = full_depth + full_missed * bg_depth
    return full_rgb, full_depth, full_missed


# ---------------------------------------------------------------------------
# Pruning (encoder.py:605-654) and splitting (geometry.py:250-274)

def prune_lattice(points, voxel_size, bits=16):
    """encoder.py:629-631: bits^3 points per voxel at offsets
    (2i+1-bits)/(bits-1) * vs/2 (corners included)."""
    return offset_points(np.asarray(points, np.float32), np.float32(voxel_size) / np.float32(2), bits=bits)


def prune_max_sigma(centers_shifted, feats, values, voxel_size, mlp, voxels, bits=16,
                    dtype=np.float64):
    """encoder.py:620-646 for the voxels listed: the bits^3 lattice around
    each (shifted) centre, the gather at that voxel, sigma from the MLP, and
    the max of relu(sigma) per voxel (the keep test is 1 - exp(-max) > th,
    encoder.py:610)."""
    voxels = np.asarray(voxels, np.int64)
    lat = prune_lattice(np.asarray(centers_shifted, np.float32)[voxels], voxel_size, bits)
    idx = np.repeat(voxels, bits ** 3)
    emb = trilinear_gather(idx, lat.reshape(-1, 3), centers_shifted, feats, values, voxel_size,
                           dtype=dtype)
    sigma, _ = mlp_forward(emb, mlp, dtype=dtype)
    return np.maximum(sigma.reshape(len(voxels), -1), 0).max(1)


def split_voxels(points, feats, values, voxel_size):
    """geometry.py:250-274: 8 children per voxel at +-vs/4, new unique sorted
    corner keys, new values by trilinear interpolation of the parent corners.
    The parent that supplies each new key is the lowest parent index among
    those sharing it (the reference's scatter_ leaves this to the device)."""
    half_voxel = np.float32(voxel_size) * np.float32(0.5)
    quarter = half_voxel * np.float32(0.5)
    pts = np.asarray(points, np.float32)
    new_points = offset_points(pts, quarter).reshape(-1, 3).astype(np.float32)
    old_coords, _ = discretize_points(pts, quarter)
    new_coords = offset_points(old_coords, 1.0).reshape(-1, 3).astype(np.int64)
    new_keys0 = offset_points(new_coords, 1.0).reshape(-1, 3).astype(np.int64)
    new_keys, inv = np.unique(new_keys0, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    parent = np.arange(new_keys0.shape[0]) // 64
    key_parent = np.full(new_keys.shape[0], np.iinfo(np.int64).max, np.int64)
    np.minimum.at(key_parent, inv, parent)
    new_feats = inv.reshape(-1, 8).astype(np.int64)
    new_values = None
    if values is not None:
        p = (new_keys - old_coords[key_parent]).astype(np.float32)[:, None, :] * np.float32(0.25) + np.float32(0.5)
        p = p[:, 0, :]
        E = np.asarray(values, np.float64)[np.asarray(feats)[key_parent]]
        out = np.zeros((new_keys.shape[0], E.shape[-1]), np.float64)
        pp = p.astype(np.float64)
        for corner in range(8):
            bx, by, bz = (corner >> 2) & 1, (corner >> 1) & 1, corner & 1
            w = (pp[:, 0] if bx else 1 - pp[:, 0]) * (pp[:, 1] if by else 1 - pp[:, 1]) * (pp[:, 2] if bz else 1 - pp[:, 2])
            out += w[:, None] * E[:, corner]
        new_values = out
    return new_points, new_feats, new_values, new_keys


# ---------------------------------------------------------------------------
# Grazing rays: hit sets that an FP64 slab test would decide differently

def aabb_intersect_f64(ray_start, ray_dir, points, voxelsize, n_max):
    """FP64 brute-force slab test (same ternaries), for counting rays whose FP32
    hit set differs because they graze a voxel face.  Vectorised over voxels,
    looped over rays: for small cases only."""
    rs = np.asarray(ray_start, np.float64).reshape(-1, 3)
    rd = np.asarray(ray_dir, np.float64).reshape(-1, 3)
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    h = float(voxelsize) * 0.5
    R = rs.shape[0]
    idx = np.full((R, n_max), -1, np.int32)
    lo_out = np.zeros((R, n_max))
    hi_out = np.zeros((R, n_max))
    with np.errstate(divide="ignore", invalid="ignore"):
        for r in range(R):
            inv = 1.0 / rd[r]
            f_low = np.zeros(pts.shape[0])
            f_high = np.full(pts.shape[0], 1e5)
            alive = np.ones(pts.shape[0], bool)
            for a in range(3):
                lo = (pts[:, a] - h - rs[r, a]) * inv[a]
                hi = (pts[:, a] + h - rs[r, a]) * inv[a]
                sw = hi < lo
                lo, hi = np.where(sw, hi, lo), np.where(sw, lo, hi)
                alive &= ~(hi < f_low)
                alive &= ~(lo > f_high)
                f_low = np.where(lo > f_low, lo, f_low)
                f_high = np.where(hi < f_high, hi, f_high)
                alive &= ~(f_low > f_high)
            ks = np.nonzero(alive)[0][:n_max]
            idx[r, :len(ks)] = ks
            lo_out[r, :len(ks)] = f_low[ks]
            hi_out[r, :len(ks)] = f_high[ks]
    return idx, lo_out, hi_out


def aabb_hits_f64(ray_start, ray_dir, points, voxelsize, n_max):
    """FP64 brute-force slab test (C++, the same comparisons in double) of
    float32 rays [R,3] against float32 box centres [N,3] of half-size
    voxelsize / 2: idx [R,n_max] i32 in point order (-1 pad), entry / exit
    depths float64, and the largest hit count of a ray."""
    rs, rd, pts = _c(ray_start, np.float32), _c(ray_dir, np.float32), _c(points, np.float32)
    R, N = rs.shape[0], pts.shape[0]
    idx = np.zeros((R, n_max), np.int32)
    lo = np.zeros((R, n_max), np.float64)
    hi = np.zeros((R, n_max), np.float64)
    most = lib().orc_aabb_intersect_f64(R, N, _p(rs), _p(rd), _p(pts),
                                        float(np.float32(voxelsize)) * 0.5, n_max, _p(idx),
                                        _p(lo), _p(hi))
    return idx, lo, hi, int(most)


def grazing_report(idx, lo, hi, ray_start, ray_dir, leaf_points, voxel_size, n_max, eps=1e-6):
    """FP32 hit rows (idx [R,P] with -1 pads; lo / hi entry and exit depths)
    of rays against the voxels with centres `leaf_points`, compared with
    FP64 slab tests of the same rays and boxes.  A ray differs when its FP32
    hit set is not the FP64 one (for a full row of P hits, when it holds a
    voxel the FP64 test misses); a difference is a graze when every voxel
    in it is crossed over less than `eps` (its FP64 interval if the FP64
    test hits it, else its FP32 one).  Also: rays with an FP32 hit shorter
    than `eps`, and the largest entry / exit depth difference over the hits
    both agree on."""
    idx = np.asarray(idx)
    R, P = idx.shape
    cap = max(4 * int(n_max), 256)
    i64, lo64, hi64, most = aabb_hits_f64(ray_start, ray_dir, leaf_points, voxel_size, cap)
    if most > cap:
        raise RuntimeError(f"grazing_report: a ray hits {most} boxes (> {cap})")
    differing = grazing = short = 0
    depth_err = 0.0
    for r in range(R):
        m32 = idx[r] != -1
        k32 = idx[r][m32]
        l32, h32 = np.asarray(lo[r])[m32], np.asarray(hi[r])[m32]
        m64 = i64[r] != -1
        pos64 = {int(v): j for j, v in enumerate(i64[r][m64])}
        s32, s64 = set(int(v) for v in k32), set(pos64)
        if ((h32 - l32) < eps).any():
            short += 1
        full = len(s32) >= n_max
        diff = (s32 - s64) if full else (s32 ^ s64)
        if diff:
            differing += 1
            pos32 = {int(v): j for j, v in enumerate(k32)}
            ok = all((hi64[r, pos64[v]] - lo64[r, pos64[v]] < eps) if v in pos64 else
                     (h32[pos32[v]] - l32[pos32[v]] < eps) for v in diff)
            grazing += int(ok)
        both = [(j, pos64[int(v)]) for j, v in enumerate(k32) if int(v) in pos64]
        if both:
            a = np.array(both)
            depth_err = max(depth_err,
                            float(np.abs(l32[a[:, 0]] - lo64[r, a[:, 1]]).max()),
                            float(np.abs(h32[a[:, 0]] - hi64[r, a[:, 1]]).max()))
    return {"rays": int(R), "eps": eps, "rays_differing_from_fp64": differing,
            "differences_that_graze": grazing, "rays_with_hits_shorter_than_eps": short,
            "max_abs_hit_depth_vs_fp64": depth_err}


# ---------------------------------------------------------------------------
# One view, end to end (fairnr_model.py:142-186 for the NSVF model with
# hierarchical sampling; nsvf.py:43-110; nerf.py:64-79; renderer.py:77-252)

def _field_eval(si, sd, sx, ray_start, ray_dir, centers_shifted, feats, values, voxel_size, mlp,
                dtype):
    """Field outputs at the valid samples of a padded [N,S] sample set, in
    ``dtype``; sample points x = o + d * z in float32 (geometry.py:160-161)."""
    valid = si != -1
    rows, cols = np.nonzero(valid)
    z = sd[rows, cols].astype(np.float32)
    x = (ray_start[rows] + (ray_dir[rows] * z[:, None]).astype(np.float32)).astype(np.float32)
    emb = trilinear_gather(si[rows, cols], x, centers_shifted, feats, values, voxel_size, dtype=dtype)
    sigma, rgb = mlp_forward(emb, mlp, dtype=dtype)
    S = np.zeros(si.shape, dtype)
    C = np.zeros(si.shape + (3,), dtype)
    S[rows, cols] = sigma
    C[rows, cols] = rgb
    return S, C


def fine_normaliser(safe):
    """Sum over the columns of `safe` [R,P] (float32) in the fixed order the
    fine sampler uses (the reference's torch sum leaves the order open): 64
    partial sums, partial j adding columns j, j + 64, ... in turn, then an
    xor butterfly over the partials, d = 1, 2, ..., 32."""
    safe = np.asarray(safe, np.float32)
    part = np.zeros((safe.shape[0], 64), np.float32)
    for k in range(safe.shape[1]):
        part[:, k % 64] = (part[:, k % 64] + safe[:, k]).astype(np.float32)
    lanes = np.arange(64)
    for d in (1, 2, 4, 8, 16, 32):
        part = (part + part[:, lanes ^ d]).astype(np.float32)
    return part[:, 0]


def prepare_fine_bins(si, sd, sx, probs, fine_steps):
    """nerf.py:64-79 (+ nsvf.py:83-87 with its ratio unset): coarse samples as
    bins [z - d/2, z + d/2], probs (p + 1e-5) / sum over all columns (fp32,
    in fine_normaliser's order), steps = fine_steps."""
    lo = (sd - (sx * np.float32(0.5)).astype(np.float32)).astype(np.float32)
    hi = (sd + (sx * np.float32(0.5)).astype(np.float32)).astype(np.float32)
    safe = (np.asarray(probs, np.float32) + np.float32(1e-5)).astype(np.float32)
    den = fine_normaliser(safe)
    p = (safe / den[:, None]).astype(np.float32)
    steps = np.full(safe.shape[0], np.float32(fine_steps), np.float32)
    return si.astype(np.int32), lo, hi, p, steps


def render_view(scene, ray_start, ray_dir, hierarchical=True, fine_steps=64.0, compat=True,
                dtype=np.float64, coarse_probs=None, use_octree=True, subset=None,
                noise_seed=None, tolerance=0.0, chunk_size=65536):
    """Render rays [R,3] of ``scene`` (a VoxelScene).  Integer and sampling
    decisions are the FP32 restatement's; gather, MLP and compositing run in
    ``dtype``.  ``coarse_probs`` ([N_hit, S_c] float32) replaces the coarse
    compositing probabilities that drive the fine sampler, so the fine
    samples can be compared with another implementation's for the same
    probabilities.

    ``subset`` (sorted ray ids): intersection and both sampling passes still
    run over every ray -- in the reference's sampler mode a ray's samples
    depend on its neighbours in the wrapper's grouping -- but the field and
    compositing only for the hit rays in the subset; the other hit rays get
    zero coarse probabilities (their fine pass then samples uniformly, which
    touches only their own outputs).  The image outputs (rgb, depth, missed)
    are then the subset's, in its order; ``eval_rows`` gives the positions
    of the evaluated rays among the hit rays.

    ``noise_seed``: training-mode sampling (encoder.py:546 with the model in
    training): both passes draw hash_noise rows of the hit rays' frame ids,
    with seeds pass_seed(noise_seed, 0) and pass_seed(noise_seed, 1); None
    samples at noise 0.5.  Returns a dict of the outputs and
    intermediates.

    ``tolerance`` > 0: the reference's early ray termination
    (``--raymarching-tolerance``) with ``chunk_size`` as its
    ``--valid-chunk-size`` in rows and samples: each pass evaluates only the
    samples march_chunks() selects (the others have field 0); the
    selections are returned as ``c_evaluated`` / ``f_evaluated`` with the
    stopped rows and their decision margins."""
    if tolerance > 0 and subset is not None:
        raise ValueError("early termination couples the rays of a row group: no subset")
    rs = np.asarray(ray_start, np.float32)
    rd = np.asarray(ray_dir, np.float32)
    vs = np.float32(scene.voxel_size)
    pts_shift = shift_centers(scene.points, vs)
    if use_octree:
        cen, ch = build_easy_octree(scene.points, vs * np.float32(0.5))
        idx, lo, hi = svo_intersect(rs[None], rd[None], cen[None], ch[None], vs, scene.max_hits)
    else:
        idx, lo, hi = aabb_intersect(rs[None], rd[None], pts_shift[None], vs, scene.max_hits)
    idx, lo, hi, hits = sort_hits(idx[0], lo[0], hi[0])
    out = {"hits": hits, "sorted_idx": idx, "sorted_lo": lo, "sorted_hi": hi}
    hit_rays = np.nonzero(hits)[0]
    if subset is None:
        ev = np.arange(len(hit_rays))
        frame_hits = hits
    else:
        subset = np.asarray(subset, np.int64)
        ev = np.nonzero(np.isin(hit_rays, subset))[0]
        frame_hits = hits[subset]
    out["eval_rows"] = ev
    idx, lo, hi = idx[hits], lo[hits], hi[hits]
    probs, steps = probs_steps(idx, lo, hi, scene.step_size)
    out.update(probs=probs, steps=steps)
    def noise(k, n_cols):
        if noise_seed is None:
            return None
        return hash_noise(pass_seed(noise_seed, k), hit_rays, n_cols)

    det = noise_seed is None
    nz = None if det else noise(0, int(np.ceil(steps).astype(np.int64).max(initial=0)) + idx.shape[1])
    si, sd, sx, S, _ = inverse_cdf_sampling(idx, lo, hi, probs, steps, -1.0, det, noise=nz,
                                            compat=compat)
    sd, sx = sample_tail(si, sd, sx)
    out.update(c_idx=si, c_depth=sd, c_dist=sx)
    rsh, rdh = rs[hits][ev], rd[hits][ev]
    sig, col = _field_eval(si[ev], sd[ev], sx[ev], rsh, rdh, pts_shift, scene.feats, scene.values,
                           vs, scene.mlp, dtype)
    if tolerance > 0:
        e, stop, margin = march_chunks(si, sig, sx, tolerance, chunk_size)
        sig, col = np.where(e, sig, 0), np.where(e[..., None], col, 0)
        out.update(c_evaluated=e, c_stopped=stop, c_margin=margin)
    rgb, depth, missed, p = composite(sig, col, sd[ev], sx[ev], si[ev], dtype=dtype)
    out.update(c_sigma=sig, c_rgb=col, c_probs=p, coarse_rgb=rgb, coarse_depth=depth,
               coarse_missed=missed)
    if hierarchical and si.shape[0] > 0:
        if coarse_probs is None:
            p_use = np.zeros(si.shape, np.float32)
            p_use[ev] = np.asarray(p, np.float32)
        else:
            p_use = np.asarray(coarse_probs, np.float32)
        fi, flo, fhi, fp, fst = prepare_fine_bins(si, sd, sx, p_use, fine_steps)
        nz = None if det else noise(1, int(math.ceil(fine_steps)) + fi.shape[1])
        si2, sd2, sx2, S2, _ = inverse_cdf_sampling(fi, flo, fhi, fp, fst, -1.0, det, noise=nz,
                                                    compat=compat)
        sd2, sx2 = sample_tail(si2, sd2, sx2)
        out.update(f_idx=si2, f_depth=sd2, f_dist=sx2)
        sig2, col2 = _field_eval(si2[ev], sd2[ev], sx2[ev], rsh, rdh, pts_shift, scene.feats,
                                 scene.values, vs, scene.mlp, dtype)
        if tolerance > 0:
            e, stop, margin = march_chunks(si2, sig2, sx2, tolerance, chunk_size)
            sig2, col2 = np.where(e, sig2, 0), np.where(e[..., None], col2, 0)
            out.update(f_evaluated=e, f_stopped=stop, f_margin=margin)
        rgb, depth, missed, p2 = composite(sig2, col2, sd2[ev], sx2[ev], si2[ev], dtype=dtype)
        out.update(f_sigma=sig2, f_probs=p2)
    full_rgb, full_depth, full_missed = fill_background(frame_hits, rgb, depth, missed, dtype=dtype)
    out.update(rgb=full_rgb, depth=full_depth, missed=full_missed)
    return out


# ---------------------------------------------------------------------------
# Training objective with the samples held fixed (finite-difference checks)

class FixedSampleLoss:
    """Sum of squared RGB errors of the hit rays for given embeddings and MLP,
    with the samples held fixed -- the part of the training loss
    (rendering_loss.py:139, RGB term) that depends on the parameters; rays
    that hit nothing contribute a constant and are left out.  Everything
    after the sample points (gather weights, MLP, compositing, background)
    is evaluated in float64; sample points are the float32 o + d * z.

    si, sd, sx: padded [N_hit, S] coarse samples of the hit rays (as
    render_view returns them); hits: boolean mask over the frame's rays."""

    def __init__(self, si, sd, sx, hits, ray_start, ray_dir, centers_shifted, feats,
                 voxel_size, target, bg=(1.0, 1.0, 1.0)):
        self.si, self.sd, self.sx = si, sd, sx
        self.frame_rows = np.nonzero(np.asarray(hits))[0]
        valid = si != -1
        self.rows, self.cols = np.nonzero(valid)
        rs = np.asarray(ray_start, np.float32)[self.frame_rows[self.rows]]
        rd = np.asarray(ray_dir, np.float32)[self.frame_rows[self.rows]]
        z = sd[self.rows, self.cols].astype(np.float32)
        x = (rs + (rd * z[:, None]).astype(np.float32)).astype(np.float32)
        vox = si[self.rows, self.cols].astype(np.int64)
        self.W = trilinear_weights(vox, x, centers_shifted, voxel_size)      # [M,8] float64
        self.keys = np.asarray(feats, np.int64)[vox]                        # [M,8]
        self.target = np.asarray(target, np.float64)[self.frame_rows]       # [N_hit,3]
        self.bg = np.asarray(bg, np.float64)

    def _select(self, rows):
        if rows is None:
            return np.arange(len(self.rows)), np.arange(self.si.shape[0])
        rows = np.unique(np.asarray(rows, np.int64))
        return np.nonzero(np.isin(self.rows, rows))[0], rows

    def embed(self, values, rows=None):
        """float64 embeddings [M_sel, D] of the samples of `rows` (all when None)."""
        m, _ = self._select(rows)
        V = np.asarray(values, np.float64)
        emb = np.zeros((len(m), V.shape[1]))
        for c in range(8):
            emb += self.W[m, c, None] * V[self.keys[m, c]]
        return emb

    def loss_from_embedding(self, emb, mlp, rows=None):
        sigma, rgb = mlp_forward(emb, mlp, dtype=np.float64)
        return self._composite_loss(sigma, rgb, rows)

    def loss(self, values, mlp, rows=None):
        return self.loss_from_embedding(self.embed(values, rows), mlp, rows)

    def _composite_loss(self, sigma, rgb, rows=None):
        m, r = self._select(rows)
        pos = np.searchsorted(r, self.rows[m])
        S = np.zeros((len(r), self.si.shape[1]))
        C = np.zeros((len(r), self.si.shape[1], 3))
        S[pos, self.cols[m]] = sigma
        C[pos, self.cols[m]] = rgb
        col, _, missed, _ = composite(S, C, self.sd[r], self.sx[r], self.si[r])
        out = col + missed[:, None] * self.bg
        return float(((out - self.target[r]) ** 2).sum())

    def mlp_central_difference(self, emb, mlp, layer, kind, pos, eps):
        """(loss(theta + eps) - loss(theta - eps)) / (2 eps) for one MLP entry
        (weights[layer][pos] if kind == "w", biases[layer][pos] if "b"),
        over all hit rays.  Exact reformulation of two loss() calls: the
        perturbation changes one output unit j of `layer`, so that layer's
        pre-activations change in column j only, and the next layer's by the
        rank-1 term outer(dh_j, W_next[:, j]); later layers are recomputed."""
        Ws = [np.asarray(w, np.float64) for w in mlp.weights]
        bs = [np.asarray(b, np.float64) for b in mlp.biases]
        n = len(Ws)
        key = (id(emb), id(mlp))
        if getattr(self, "_base_key", None) != key:     # unperturbed layers, kept between calls
            pre, act = [], []
            h = emb
            for i in range(n):
                a = h @ Ws[i].T + bs[i]
                pre.append(a)
                h = np.maximum(a, 0) if i < n - 1 else a
                act.append(h)
            self._base_key, self._base = key, (pre, act)
        pre, act = self._base
        j = pos[0]
        col_in = emb if layer == 0 else act[layer - 1]
        du = col_in[:, pos[1]] if kind == "w" else np.ones(emb.shape[0])
        vals = []
        for delta in (eps, -eps):
            aj = pre[layer][:, j] + delta * du
            if layer == n - 1:
                out = pre[layer].copy()
                out[:, j] = aj
            else:
                dh = np.maximum(aj, 0) - act[layer][:, j]
                a = pre[layer + 1] + np.outer(dh, Ws[layer + 1][:, j])
                for i in range(layer + 1, n):
                    if i > layer + 1:
                        a = h @ Ws[i].T + bs[i]
                    h = np.maximum(a, 0) if i < n - 1 else a
                out = h
            vals.append(self._composite_loss(out[:, 0], 1.0 / (1.0 + np.exp(-out[:, 1:4]))))
        return (vals[0] - vals[1]) / (2 * eps)

    def rows_touching_key(self, key):
        """Hit rays with a sample in a voxel that has corner `key`."""
        m = np.nonzero((self.keys == key).any(1))[0]
        return np.unique(self.rows[m])
