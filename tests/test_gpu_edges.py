"""Edge cases of the fused renderer against the oracle: a view that hits
nothing, a one-voxel scene (the octree's root is the leaf), rays that start
inside voxels (entry depth 0), ray counts that are not a multiple of a wave,
and a one-hit row budget."""
import dataclasses

import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell
from nsvf_amd import scene_io as IO
from nsvf_amd.render import DeviceScene, Renderer

pytestmark = pytest.mark.gpu

BG = np.float32(1.0)


def N(t):
    return t.detach().cpu().numpy()


def render(sc, rs, rd, device, precision="fp32", width=0, keep=True):
    r = Renderer(device, precision=precision, hierarchical=True, compat=False)
    out = r.render(DeviceScene(sc, device), torch.from_numpy(np.ascontiguousarray(rs)).to(device),
                   torch.from_numpy(np.ascontiguousarray(rd)).to(device), keep=keep, width=width)
    return r, [N(t) for t in out]


def check_hits(r, sc, rs, rd):
    ref = O.render_view(sc, rs, rd, compat=False)
    k = r.kept
    assert np.array_equal(N(k["flag"]).astype(bool), ref["hits"])
    cnt, rows = N(k["count"]), N(k["hits"])
    assert np.array_equal(cnt, (ref["sorted_idx"] != -1).sum(1))
    for ray in np.nonzero(ref["hits"])[0]:
        n = cnt[ray]
        assert np.array_equal(rows[ray, :n, 2], ref["sorted_idx"][ray, :n])
        assert np.array_equal(rows[ray, :n, 0].view(np.float32), ref["sorted_lo"][ray, :n])
    return ref


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_view_that_hits_nothing(device, precision):
    """Rays pointing away from the scene: no hit rays, the background."""
    sc = shell()
    rng = np.random.default_rng(0)
    rs = np.tile(np.array([[0.0, 0.0, 3.0]], np.float32), (100, 1))
    rd = rng.standard_normal((100, 3)).astype(np.float32)
    rd[:, 2] = np.abs(rd[:, 2]) + 0.5
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    for width in (0, 10):
        r, (rgb, depth, missed) = render(sc, rs, rd, device, precision, width=width)
        assert r.last.hit_rays == 0
        assert np.array_equal(missed, np.ones(100, np.float32))
        assert np.array_equal(rgb, np.full((100, 3), BG, np.float32))


def test_one_voxel_scene(device):
    """A single voxel: the octree is one leaf, the traversal's root record
    holds it; hits and image against the oracle."""
    sc = IO.scene_from_points(np.array([[0.1, -0.2, 0.05]], np.float32), 0.25, embed_dim=32)
    rng = np.random.default_rng(1)
    tgt = sc.points[0] + rng.uniform(-0.1, 0.1, (77, 3)).astype(np.float32)
    rs = np.tile(np.array([[1.5, 0.7, -0.9]], np.float32), (77, 1))
    rd = (tgt - rs).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    r, (rgb, depth, missed) = render(sc, rs, rd, device)
    ref = check_hits(r, sc, rs, rd)
    assert ref["hits"].all()
    assert np.abs(rgb - ref["rgb"]).max() < 1e-5 and np.abs(missed - ref["missed"]).max() < 1e-5


def test_rays_starting_inside_voxels(device):
    """Origins at voxel centres: entry depth 0 on the first hit."""
    sc = shell()
    rng = np.random.default_rng(2)
    pick = rng.choice(sc.num_voxels, 130, replace=False)
    rs = sc.points[pick].astype(np.float32)
    rd = rng.standard_normal((130, 3)).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    r, (rgb, depth, missed) = render(sc, rs, rd, device)
    ref = check_hits(r, sc, rs, rd)
    assert (ref["sorted_lo"][:, 0] == 0).sum() > 100
    assert np.abs(rgb - ref["rgb"]).max() < 1e-5


@pytest.mark.parametrize("R", [1, 63, 65])
def test_ray_counts_off_the_wave(device, R):
    sc = shell()
    rng = np.random.default_rng(3 + R)
    cam = np.array([2.0, 1.5, -0.8], np.float32)
    rs = np.tile(cam, (R, 1))
    tgt = rng.uniform(-0.5, 0.5, (R, 3)).astype(np.float32)
    rd = (tgt - rs).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    r, (rgb, depth, missed) = render(sc, rs, rd, device)
    ref = check_hits(r, sc, rs, rd)
    assert np.abs(rgb - ref["rgb"]).max() < 1e-5


def test_one_hit_budget(device):
    """max_hits = 1: every hit row is cut to the reference traversal's first
    leaf, then sorted (one element)."""
    sc = dataclasses.replace(shell(), max_hits=1)
    rng = np.random.default_rng(4)
    rs = np.tile(np.array([0.3, -2.2, 0.9], np.float32), (200, 1))
    tgt = rng.uniform(-0.7, 0.7, (200, 3)).astype(np.float32)
    rd = (tgt - rs).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    r, (rgb, depth, missed) = render(sc, rs, rd, device)
    ref = check_hits(r, sc, rs, rd)
    assert ref["hits"].sum() > 50
    assert np.abs(rgb - ref["rgb"]).max() < 1e-5


def padded(cnt, off, arr, width, fill):
    out = np.full((len(cnt), width), fill, dtype=arr.dtype)
    for n in range(len(cnt)):
        out[n, : cnt[n]] = arr[off[n]: off[n] + cnt[n]]
    return out


@pytest.fixture(scope="module")
def cube():
    """A dense 56^3 cube of voxels and rays through it, many along the
    diagonals: rows of more than 128 hits (past the coarse sampler's staged
    bins) and of more than 384 coarse samples (past the fine sampler's)."""
    n = 56
    g = (np.arange(n) - (n - 1) / 2) * 0.1
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    sc = dataclasses.replace(IO.scene_from_points(pts, 0.1, embed_dim=32), max_hits=200,
                             step_size=np.float32(0.025))
    rng = np.random.default_rng(5)
    R = 96
    d = rng.normal(size=(R, 3)).astype(np.float32)
    d[: R // 2] = np.sign(d[: R // 2]) * (1 + 0.05 * rng.random((R // 2, 3)))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return sc, (-d * 3.0).astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("compat", [False, True])
def test_rows_past_the_staged_bins(device, cube, compat):
    """Rows too long for the samplers' LDS staging take the sequential walk;
    hit rows, coarse and fine streams still equal the oracle's bit for bit
    and the FP32 image is within 1e-5 of the FP64 one."""
    sc, rs, rd = cube
    r = Renderer(device, precision="fp32", hierarchical=True, compat=compat)
    rgb, depth, missed = (N(t) for t in r.render(
        DeviceScene(sc, device), torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device),
        keep=True))
    k = r.kept
    ref = O.render_view(sc, rs, rd, compat=compat)
    hit_cnt = (ref["sorted_idx"] != -1).sum(1)
    assert (hit_cnt > 128).sum() > 10
    cnt, rows = N(k["count"]), N(k["hits"])
    assert np.array_equal(cnt, hit_cnt)
    for ray in np.nonzero(ref["hits"])[0]:
        m = cnt[ray]
        assert np.array_equal(rows[ray, :m, 2], ref["sorted_idx"][ray, :m])
        assert np.array_equal(rows[ray, :m, 0].view(np.float32), ref["sorted_lo"][ray, :m])
    c_cnt, c_off = N(k["c_cnt"]), N(k["c_off"])
    ci = ref["c_idx"]
    assert (c_cnt > 384).sum() > 10
    W = ci.shape[1]
    assert np.array_equal(padded(c_cnt, c_off, N(k["c_samples"][0]), W, -1), ci)
    assert np.array_equal(padded(c_cnt, c_off, N(k["c_samples"][1]), W,
                                 np.float32(O.MAX_DEPTH)), ref["c_depth"])
    gp = padded(c_cnt, c_off, N(k["c_prob"])[: c_off[-1] + c_cnt[-1]], W, np.float32(0))
    ref2 = O.render_view(sc, rs, rd, compat=compat, coarse_probs=gp)
    f_cnt, f_off = N(k["f_cnt"]), N(k["f_off"])
    fi = ref2["f_idx"]
    W2 = fi.shape[1]
    assert np.array_equal(padded(f_cnt, f_off, N(k["f_samples"][0]), W2, -1), fi)
    assert np.array_equal(padded(f_cnt, f_off, N(k["f_samples"][1]), W2,
                                 np.float32(O.MAX_DEPTH)), ref2["f_depth"])
    assert np.abs(rgb - ref2["rgb"]).max() < 1e-5
    assert np.abs(missed - ref2["missed"]).max() < 1e-5
