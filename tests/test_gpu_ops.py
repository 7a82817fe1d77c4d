"""Parity of the HIP op kernels (through the C ABI) with the CPU oracle.

Bar: indices, counts and padded layouts bit-exact; depths and distances
bit-exact as well (both sides use the same FP32 operation order and IEEE
reciprocal/division), with any mismatch reported.
"""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import axis_rays, hits_for_sampling, random_rays, shell, shell_octree, shell_rays
from nsvf_amd import _ext, clib

pytestmark = pytest.mark.gpu


def T(a, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def assert_same(gpu, ref, what):
    g, r = N(gpu), np.asarray(ref)
    assert g.shape == r.shape, f"{what}: shape {g.shape} vs {r.shape}"
    if g.dtype.kind == "f":
        bad = ~((g == r) | (np.isnan(g) & np.isnan(r)))
    else:
        bad = g != r
    if bad.any():
        i = np.argwhere(bad)[:5]
        raise AssertionError(f"{what}: {bad.sum()} mismatches, first at {i.tolist()}: "
                             f"gpu {g[tuple(i[0])]} ref {r[tuple(i[0])]}")


def _aabb_case(dev, rs, rd, n_max, sc=None):
    sc = sc or shell()
    pts = O.shift_centers(sc.points, sc.voxel_size)
    ref = O.aabb_intersect(rs[None], rd[None], pts[None], sc.voxel_size, n_max)
    out = _ext.aabb_intersect(T(rs[None], dev), T(rd[None], dev), T(pts[None], dev),
                              sc.voxel_size, n_max)
    torch.cuda.synchronize()
    for o, r, w in zip(out, ref, ("idx", "min_depth", "max_depth")):
        assert_same(o, r, f"aabb {w}")
    return ref


@pytest.mark.parametrize("n_max", [60, 5])
def test_aabb_camera(device, n_max):
    rs, rd = shell_rays()
    ref = _aabb_case(device, rs, rd, n_max)
    assert (ref[0] != -1).any(-1).sum() > 1000


def test_aabb_random_and_axis_rays(device):
    rs, rd = random_rays(3000, seed=3)
    _aabb_case(device, rs, rd, 60)
    rs, rd = axis_rays(shell(), 2000, seed=4)
    _aabb_case(device, rs, rd, 60)


def test_aabb_batched_unshared_points(device):
    sc = shell()
    pts = O.shift_centers(sc.points, sc.voxel_size)
    rng = np.random.default_rng(5)
    B = 3
    rs, rd = random_rays(B * 700, seed=6)
    rs, rd = rs.reshape(B, 700, 3), rd.reshape(B, 700, 3)
    # each batch sees a differently translated copy of the scene
    P = np.stack([pts + rng.uniform(-0.1, 0.1, 3).astype(np.float32) for _ in range(B)])
    ref = O.aabb_intersect(rs, rd, P, sc.voxel_size, 60)
    out = _ext.aabb_intersect(T(rs, device), T(rd, device), T(P, device), sc.voxel_size, 60)
    for o, r in zip(out, ref):
        assert_same(o, r, "aabb batched")


def test_aabb_empty(device):
    sc = shell()
    pts = T(O.shift_centers(sc.points, sc.voxel_size)[None], device)
    z = torch.zeros((1, 0, 3), device=device)
    idx, lo, hi = _ext.aabb_intersect(z, z, pts, sc.voxel_size, 60)
    assert idx.shape == (1, 0, 60)


def _svo_case(dev, rs, rd, n_max):
    sc = shell()
    cen, ch = shell_octree()
    ref = O.svo_intersect(rs[None], rd[None], cen[None], ch[None], sc.voxel_size, n_max)
    out = _ext.svo_intersect(T(rs[None], dev), T(rd[None], dev), T(cen[None], dev),
                             T(ch[None], dev), sc.voxel_size, n_max)
    torch.cuda.synchronize()
    for o, r, w in zip(out, ref, ("idx", "min_depth", "max_depth")):
        assert_same(o, r, f"svo {w}")
    return ref


@pytest.mark.parametrize("n_max", [60, 7])
def test_svo_camera(device, n_max):
    rs, rd = shell_rays()
    _svo_case(device, rs, rd, n_max)


def test_svo_random_axis(device):
    rs, rd = random_rays(3000, seed=8)
    _svo_case(device, rs, rd, 60)
    rs, rd = axis_rays(shell(), 2000, seed=9)
    _svo_case(device, rs, rd, 60)


def _icdf_inputs(seed=0, n_max=None):
    sc = shell()
    rs, rd = shell_rays(seed=7 + seed)
    return hits_for_sampling(sc, rs, rd, n_max=n_max)


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("n_max", [60, 6])
def test_icdf_op(device, compat, n_max):
    idx, lo, hi, probs, steps, _ = _icdf_inputs(0, n_max)
    Nr, P = idx.shape
    S = int(np.ceil(steps).max()) + P
    rng = np.random.default_rng(11)
    noise = rng.uniform(0.001, 0.999, (1, Nr, S)).astype(np.float32)
    ref = O.inverse_cdf_sampling_op(idx[None], lo[None], hi[None], noise, probs[None],
                                    steps[None], -1.0, compat)
    out = _ext.inverse_cdf_sampling(T(idx[None], device), T(lo[None], device), T(hi[None], device),
                                    T(noise, device), T(probs[None], device), T(steps[None], device),
                                    -1.0, compat=compat)
    for o, r, w in zip(out, ref[:3], ("idx", "depth", "dists")):
        assert_same(o, r, f"icdf {w}")


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("n_max", [60, 6])
@pytest.mark.parametrize("deterministic", [True, False])
def test_icdf_grouped(device, compat, n_max, deterministic):
    """The one-launch row kernel against the oracle's restatement of the
    reference wrapper (G=200 regrouping, 800-column chunks)."""
    idx, lo, hi, probs, steps, _ = _icdf_inputs(1, n_max)
    Nr, P = idx.shape
    S = int(np.ceil(steps).max()) + P
    noise = None
    if not deterministic:
        noise = np.random.default_rng(12).uniform(0.001, 0.999, (Nr, S)).astype(np.float32)
    ri, rd_, rx, rS, _ = O.inverse_cdf_sampling(idx, lo, hi, probs, steps, -1.0,
                                                deterministic=deterministic, noise=noise,
                                                compat=compat)
    oi, od, ox, S2 = clib.inverse_cdf_sampling_rows(
        T(idx, device), T(lo, device), T(hi, device), T(probs, device), T(steps, device), -1.0,
        None if noise is None else T(noise, device), compat=compat)
    assert S2 == rS
    L = ri.shape[1]
    assert_same(oi[:, :L], ri, "icdf rows idx")
    assert_same(od[:, :L], rd_, "icdf rows depth")
    assert_same(ox[:, :L], rx, "icdf rows dists")
    assert (N(oi[:, L:]) == -1).all()


def test_icdf_grouped_many_rows(device):
    """More than 200 x 800 rows, so several chunks and the chunk-edge
    neighbours are exercised; rows are tiled copies with shuffled order."""
    idx, lo, hi, probs, steps, _ = _icdf_inputs(2, 8)
    rng = np.random.default_rng(13)
    sel = rng.integers(0, idx.shape[0], 170_001)
    idx, lo, hi, probs, steps = idx[sel], lo[sel], hi[sel], probs[sel], steps[sel]
    ri, rd_, rx, rS, _ = O.inverse_cdf_sampling(idx, lo, hi, probs, steps, -1.0, True, compat=True)
    oi, od, ox, S2 = clib.inverse_cdf_sampling_rows(
        T(idx, device), T(lo, device), T(hi, device), T(probs, device), T(steps, device), -1.0,
        None, compat=True)
    L = ri.shape[1]
    assert_same(oi[:, :L], ri, "icdf many idx")
    assert_same(od[:, :L], rd_, "icdf many depth")
    assert_same(ox[:, :L], rx, "icdf many dists")


@pytest.mark.parametrize("deterministic", [True, False])
def test_uniform_op(device, deterministic):
    sc = shell()
    idx, lo, hi, probs, steps, _ = _icdf_inputs(3)
    Nr, P = idx.shape
    lengths = np.where(idx != -1, hi, 0).max(1) - lo[:, 0]
    S = int(lengths.max() / sc.step_size) + 2 * P
    if deterministic:
        noise = np.full((1, Nr, S), 0.5, np.float32)
    else:
        noise = np.random.default_rng(14).uniform(0.0, 1.0, (1, Nr, S)).astype(np.float32)
    ref = O.uniform_ray_sampling_op(idx[None], lo[None], hi[None], noise, sc.step_size, S)
    out = _ext.uniform_ray_sampling(T(idx[None], device), T(lo[None], device), T(hi[None], device),
                                    T(noise, device), sc.step_size, S)
    for o, r, w in zip(out, ref[:3], ("idx", "depth", "dists")):
        assert_same(o, r, f"uniform {w}")


def test_clib_wrappers_match_oracle(device):
    sc = shell()
    rs, rd = shell_rays(seed=21)
    pts = O.shift_centers(sc.points, sc.voxel_size)
    cen, ch = shell_octree()
    a = clib.aabb_ray_intersect(sc.voxel_size, sc.max_hits, T(pts[None], device),
                                T(rs[None], device), T(rd[None], device))
    b = clib.svo_ray_intersect(sc.voxel_size, sc.max_hits, T(cen[None], device),
                               T(ch[None], device), T(rs[None], device), T(rd[None], device))
    ra = O.aabb_ray_intersect(sc.voxel_size, sc.max_hits, pts[None], rs[None], rd[None])
    rb = O.svo_ray_intersect(sc.voxel_size, sc.max_hits, cen[None], ch[None], rs[None], rd[None])
    for o, r in zip(a, ra):
        assert_same(o, r, "clib aabb")
    for o, r in zip(b, rb):
        assert_same(o, r, "clib svo")


def test_errors_raise(device):
    sc = shell()
    pts = torch.zeros((1, 4, 3))
    with pytest.raises(RuntimeError):
        _ext.aabb_intersect(torch.zeros(1, 2, 3), torch.zeros(1, 2, 3), pts, 0.1, 4)
    with pytest.raises(RuntimeError):
        _ext.aabb_intersect(torch.zeros(1, 2, 3, device=device),
                            torch.zeros(1, 2, 3, device=device).double(), pts.to(device), 0.1, 4)
