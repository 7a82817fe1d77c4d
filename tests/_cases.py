"""Seeded inputs shared by the CPU and GPU tests (NumPy only)."""
from __future__ import annotations

import functools

import numpy as np

import nsvf_oracle as O
from nsvf_amd import scene as S


@functools.lru_cache(maxsize=None)
def shell():
    return S.shell_scene()


@functools.lru_cache(maxsize=None)
def shell_octree():
    sc = shell()
    return O.build_easy_octree(sc.points, np.float32(sc.voxel_size) * np.float32(0.5))


def shell_rays(width=64, seed=7, radius=2.5):
    cam = S.sphere_camera(radius, seed)
    return S.camera_rays(width, width, cam)


def random_rays(n, seed, radius=2.0, spread=0.6):
    """Rays from random points on a sphere aimed near the origin."""
    rng = np.random.default_rng(seed)
    o = rng.standard_normal((n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * radius
    tgt = rng.uniform(-spread, spread, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def axis_rays(sc, n, seed):
    """Axis-aligned rays (two zero direction components), some with origins
    exactly on voxel face planes (NaN slab bounds), some starting inside the
    grid (entry clipped to 0)."""
    rng = np.random.default_rng(seed)
    vs = np.float32(sc.voxel_size)
    o = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32)
    d[np.arange(n), ax] = sign
    o[np.arange(n), ax] = -sign * np.float32(1.5)
    # snap a third of the rays' other coordinates onto face planes k*vs - 1
    snap = rng.random(n) < 0.33
    k = rng.integers(0, 33, (n, 3)).astype(np.float32)
    planes = (np.float32(-1.0) + k * vs).astype(np.float32)
    for a in range(3):
        m = snap & (ax != a)
        o[m, a] = planes[m, a]
    inside = rng.random(n) < 0.2
    o[inside] = rng.uniform(-0.9, 0.9, (inside.sum(), 3)).astype(np.float32)
    return o, d


def hits_for_sampling(sc, rs, rd, n_max=None, use_svo=False):
    """Sorted hits of the hit rays, probs and steps (the sampler's inputs)."""
    n_max = sc.max_hits if n_max is None else n_max
    if use_svo:
        cen, ch = shell_octree()
        idx, lo, hi = O.svo_intersect(rs[None], rd[None], cen[None], ch[None], sc.voxel_size, n_max)
    else:
        pts = O.shift_centers(sc.points, sc.voxel_size)
        idx, lo, hi = O.aabb_intersect(rs[None], rd[None], pts[None], sc.voxel_size, n_max)
    idx, lo, hi, hits = O.sort_hits(idx[0], lo[0], hi[0])
    idx, lo, hi = idx[hits], lo[hits], hi[hits]
    probs, steps = O.probs_steps(idx, lo, hi, sc.step_size)
    return idx, lo, hi, probs, steps, hits
