"""Synthetic sparse-voxel scenes and cameras for the measured workloads.

A scene holds the reference encoder's voxel state (fairnr/modules/encoder.py:
285-310): voxel centres, corner keys, each voxel's 8 corner ids, the corner
embedding table, voxel size, step size and max hits; plus the field MLP's
weights.  Everything is generated from fixed seeds in NumPy (float32), so the
same inputs reach the HIP path and any CPU check.

Workloads:
  shell      32^3 grid over [-1,1]^3 (voxel 0.0625), voxels whose centre
             radius is in [0.5, 0.8]: 6,568 voxels, 8,796 corner keys;
             32-d embeddings ~ N(0, 0.1^2); MLP 3 x 128 ReLU.
  noise      128^3 grid over [-1,1]^3 (voxel 2^-6), the cells where seeded
             3-octave value noise is highest, 100,000 voxels by default.
Cameras: pinhole on a sphere around the origin looking at it, focal length
1.2 x image width, integer pixel coordinates (fairnr/data/data_utils.py:57-65).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class FieldMLP:
    """L hidden ReLU layers then a 4-wide head: sigma = out[0],
    rgb = sigmoid(out[1:4]).  weights[i] is [out, in]."""
    weights: list
    biases: list

    @property
    def dims(self):
        return [self.weights[0].shape[1]] + [w.shape[0] for w in self.weights]

    def flops_per_sample(self):
        return 2 * sum(w.shape[0] * w.shape[1] for w in self.weights)


def make_mlp(in_dim, hidden, layers, seed, head_gain=20.0, sigma_bias=0.5):
    """Kaiming-normal weights (fan_in, ReLU gain, as implicit.py:73-75) and
    nn.Linear-style uniform biases, from a fixed seed.  The head is scaled by
    ``head_gain`` and the sigma bias set to ``sigma_bias`` so that a random
    field over N(0, 0.1^2) embeddings renders a partly opaque, coloured image
    (unscaled, sigma stays within +-0.15 and every view is nearly empty)."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    d = in_dim
    for i in range(layers + 1):
        out = hidden if i < layers else 4
        ws.append((rng.standard_normal((out, d)) * math.sqrt(2.0 / d)).astype(np.float32))
        bound = 1.0 / math.sqrt(d)
        bs.append(rng.uniform(-bound, bound, out).astype(np.float32))
        d = out
    ws[-1] = (ws[-1] * np.float32(head_gain)).astype(np.float32)
    bs[-1][0] = np.float32(sigma_bias)
    return FieldMLP(ws, bs)


@dataclass
class VoxelScene:
    points: np.ndarray        # [N,3] f32 voxel centres
    voxel_size: float
    step_size: float
    max_hits: int
    embed_dim: int
    keys: np.ndarray = None   # [K,3] i64 corner keys (half-voxel units)
    feats: np.ndarray = None  # [N,8] i64 corner ids per voxel
    values: np.ndarray = None  # [K,D] f32 corner embeddings
    mlp: FieldMLP = None
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def num_voxels(self):
        return int(self.points.shape[0])

    @property
    def num_keys(self):
        return int(self.keys.shape[0])


def _corner_keys(points, voxel_size):
    """encoder.py:269-276 in NumPy: corner keys unique and sorted, and each
    voxel's 8 corner ids in offset order 4x + 2y + z."""
    vs = np.float32(voxel_size)
    half = vs * np.float32(0.5)
    mn = points.min(0, keepdims=True)
    coords = np.round((points - mn) / half).astype(np.int64)
    off = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.int64)
    keys0 = (coords[:, None, :] + off[None]).reshape(-1, 3)
    keys, inv = np.unique(keys0, axis=0, return_inverse=True)
    return keys, inv.reshape(-1, 8).astype(np.int64)


def _finish(points, voxel_size, step_ratio, max_hits, embed_dim, emb_seed, mlp, name):
    keys, feats = _corner_keys(points, voxel_size)
    rng = np.random.default_rng(emb_seed)
    values = (rng.standard_normal((keys.shape[0], embed_dim)) * 0.1).astype(np.float32)
    return VoxelScene(points=points.astype(np.float32), voxel_size=float(np.float32(voxel_size)),
                      step_size=float(np.float32(voxel_size * step_ratio)), max_hits=max_hits,
                      embed_dim=embed_dim, keys=keys, feats=feats, values=values, mlp=mlp,
                      name=name)


def grid_centres(res, lo=-1.0, hi=1.0):
    vs = np.float32((hi - lo) / res)
    c = (np.float32(lo) + vs * (np.arange(res, dtype=np.float32) + np.float32(0.5))).astype(np.float32)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([X, Y, Z], -1).reshape(-1, 3), float(vs)


def shell_scene(res=32, r_in=0.5, r_out=0.8, embed_dim=32, hidden=128, layers=3,
                emb_seed=0, mlp_seed=1, max_hits=60, step_ratio=0.125, jitter_seed=None):
    """Spherical-shell scene (the CPU-runnable workload).  With
    ``jitter_seed`` the shell radii get a seeded +-5% perturbation, giving
    differently sized scenes from the same generator."""
    P, vs = grid_centres(res)
    if jitter_seed is not None:
        rng = np.random.default_rng(jitter_seed)
        r_in *= 1 + rng.uniform(-0.05, 0.05)
        r_out *= 1 + rng.uniform(-0.05, 0.05)
    r = np.linalg.norm(P.astype(np.float64), axis=1)
    pts = P[(r >= r_in) & (r <= r_out)]
    mlp = make_mlp(embed_dim, hidden, layers, mlp_seed)
    return _finish(pts, vs, step_ratio, max_hits, embed_dim, emb_seed, mlp, f"shell{res}")


def value_noise(res, octaves=3, base_cells=8, seed=0):
    """Seeded value noise on a res^3 grid: per octave, random values on a
    coarse lattice (base_cells * 2^o cells per side), trilinearly
    interpolated; octave amplitudes 1, 1/2, 1/4."""
    rng = np.random.default_rng(seed)
    out = np.zeros((res, res, res), np.float64)
    t = (np.arange(res) + 0.5) / res
    for o in range(octaves):
        cells = base_cells * (2 ** o)
        lat = rng.random((cells + 1, cells + 1, cells + 1))
        u = t * cells
        i0 = np.minimum(np.floor(u).astype(np.int64), cells - 1)
        f = u - i0
        f = f * f * (3 - 2 * f)
        fx, fy, fz = np.meshgrid(f, f, f, indexing="ij")
        ix, iy, iz = np.meshgrid(i0, i0, i0, indexing="ij")
        acc = np.zeros_like(out)
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                    acc += w * lat[ix + dx, iy + dy, iz + dz]
        out += acc * (0.5 ** o)
    return out


def noise_scene(res=128, num_voxels=100_000, seed=0, embed_dim=32, hidden=128, layers=3,
                emb_seed=0, mlp_seed=1, max_hits=60, step_ratio=0.125):
    """Value-noise scene: the ``num_voxels`` cells of highest noise value."""
    P, vs = grid_centres(res)
    v = value_noise(res, seed=seed).reshape(-1)
    order = np.argsort(-v, kind="stable")[:num_voxels]
    pts = P[np.sort(order)]
    mlp = make_mlp(embed_dim, hidden, layers, mlp_seed)
    sc = _finish(pts, vs, step_ratio, max_hits, embed_dim, emb_seed, mlp, f"noise{res}")
    sc.meta["threshold"] = float(v[order[-1]])
    return sc


# ---------------------------------------------------------------------------
# Cameras

def look_at_c2w(cam_pos, at=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0)):
    """Camera-to-world rotation (columns x, y, z) of a camera at ``cam_pos``
    looking at ``at`` (fairnr/data/geometry.py:125-160, OpenCV convention)."""
    cam = np.asarray(cam_pos, np.float64)
    z = np.asarray(at, np.float64) - cam
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    y /= np.linalg.norm(y)
    return np.stack([x, y, z], 1)


def camera_rays(width, height, cam_pos, focal=None):
    """Rays of a pinhole camera at integer pixel coordinates (u = column,
    v = row; data_utils.py:57-65, geometry.py:76-122): direction =
    normalise(R [(u-cx)/f, (v-cy)/f, 1]).  Computed in float64, returned as
    float32 ray_start [R,3] and ray_dir [R,3] in row-major pixel order."""
    f = 1.2 * width if focal is None else focal
    cx, cy = width / 2.0, height / 2.0
    R = look_at_c2w(cam_pos)
    v, u = np.mgrid[0:height, 0:width].astype(np.float64)
    d_cam = np.stack([(u - cx) / f, (v - cy) / f, np.ones_like(u)], -1).reshape(-1, 3)
    d = d_cam @ R.T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    start = np.broadcast_to(np.asarray(cam_pos, np.float64), d.shape)
    return start.astype(np.float32).copy(), d.astype(np.float32)


def sphere_camera(radius, seed):
    """A camera position at ``radius`` in a seeded generic direction."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(3)
    v /= np.linalg.norm(v)
    return v * radius


def to_device(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)
