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
