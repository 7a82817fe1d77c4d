"""Pruning and splitting on the GPU against the oracle (spherical-shell scene).

Bars: split child centres, corner keys and key ids bit-exact; split
embeddings within 1e-6 of the FP64 interpolation; FP32 pruning scores within
1e-4 (relative to max(1, |score|)) of the FP64 lattice evaluation and the
same keep decisions wherever the score is farther than that from the
threshold; BF16 scores within 5e-2."""
import math

import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell, shell_rays
from nsvf_amd.render import DeviceScene, Renderer
from nsvf_amd.voxels import prune_scene, prune_scores, pruning, split_scene

pytestmark = pytest.mark.gpu

LN2 = math.log(2.0)


@pytest.fixture(scope="module")
def dscene(device):
    sc = shell()
    return sc, DeviceScene(sc, device)


def test_prune_scores_match_oracle(dscene):
    sc, dsc = dscene
    s32 = prune_scores(dsc, "fp32").cpu().numpy()
    s16 = prune_scores(dsc, "bf16").cpu().numpy()
    assert s32.shape == (sc.num_voxels,) and (s32 >= 0).all()
    rng = np.random.default_rng(3)
    sel = np.sort(rng.choice(sc.num_voxels, 96, replace=False))
    ref = O.prune_max_sigma(O.shift_centers(sc.points, sc.voxel_size), sc.feats, sc.values,
                            sc.voxel_size, sc.mlp, sel)
    scale = np.maximum(1.0, np.abs(ref))
    assert np.abs(s32[sel] - ref).max() / 1.0 < 1e-4 * scale.max()
    assert (np.abs(s32[sel] - ref) <= 1e-4 * scale).all()
    assert (np.abs(s16[sel] - ref) <= 5e-2 * scale).all()
    keep_ref = (1 - np.exp(-ref)) > 0.5
    keep = (1 - np.exp(-s32[sel].astype(np.float64))) > 0.5
    clear = np.abs(ref - LN2) > 1e-4 * scale
    assert np.array_equal(keep[clear], keep_ref[clear])


def test_pruning_keeps_both_kinds_and_scene_renders(dscene, device):
    sc, dsc = dscene
    keep = pruning(dsc, th=0.5)
    k = keep.cpu().numpy()
    assert 0 < k.sum() < sc.num_voxels, "the shell field should prune some voxels and keep some"
    pruned = prune_scene(sc, keep)
    assert pruned.num_voxels == int(k.sum()) and pruned.num_keys == sc.num_keys
    # a pruned scene renders; rays that hit only pruned voxels now miss
    dp = DeviceScene(pruned, device)
    rs, rd = shell_rays(width=32, seed=9)
    r = Renderer(device, precision="fp32", hierarchical=False)
    t = lambda a: torch.from_numpy(a).to(device)
    rgb, depth, missed = r.render(dp, t(rs), t(rd))
    torch.cuda.synchronize()
    ref = O.render_view(pruned, rs, rd, hierarchical=False)
    assert np.array_equal(r.coarse_stream(dp, t(rs), t(rd)).flag.cpu().numpy().astype(bool)[:rs.shape[0]],
                          ref["hits"])
    assert np.abs(rgb.cpu().numpy() - ref["rgb"]).max() < 1e-4


def test_split_matches_oracle(device):
    sc = shell()
    new = split_scene(sc, device)
    pts_shift = O.shift_centers(sc.points, sc.voxel_size)
    rp, rf, rv, rk = O.split_voxels(pts_shift, sc.feats, sc.values, sc.voxel_size)
    assert np.array_equal(new.points, rp.astype(np.float32))
    assert np.array_equal(new.keys, rk)
    assert np.array_equal(new.feats, rf)
    assert np.abs(new.values - rv).max() < 1e-6
    assert new.voxel_size == float(np.float32(sc.voxel_size) * np.float32(0.5))
    assert new.max_hits == sc.max_hits * 1.5
    assert new.num_voxels == 8 * sc.num_voxels


def test_split_small_and_isolated_voxels(device):
    """Two far-apart voxels share no corners: 2 x 27 keys; one voxel: 27."""
    sc = shell()
    for pts in (sc.points[:1], sc.points[[0, 3000]]):
        keys, feats = O.voxel_keys(pts, sc.voxel_size)
        rng = np.random.default_rng(1)
        vals = rng.standard_normal((len(keys), 8)).astype(np.float32)
        small = type(sc)(points=pts.astype(np.float32), voxel_size=sc.voxel_size,
                         step_size=sc.step_size, max_hits=sc.max_hits, embed_dim=8, keys=keys,
                         feats=feats, values=vals, mlp=sc.mlp)
        new = split_scene(small, device)
        assert new.num_keys == 27 * len(pts)
        rp, rf, rv, rk = O.split_voxels(O.shift_centers(pts, sc.voxel_size), feats, vals,
                                        sc.voxel_size)
        assert np.array_equal(new.keys, rk) and np.array_equal(new.feats, rf)
        assert np.abs(new.values - rv).max() < 1e-6


def test_pruning_over_an_interval(dscene):
    """delta: keep where 1 - exp(-max relu(sigma) * delta) > th; delta = 1 is
    the reference's rule, a short interval keeps only the densest voxels."""
    sc, dsc = dscene
    s = prune_scores(dsc, "fp32")
    assert torch.equal(pruning(dsc, th=0.5, delta=1.0), pruning(dsc, th=0.5))
    step = sc.step_size
    k = pruning(dsc, th=0.5, delta=step)
    assert torch.equal(k, (1 - torch.exp(-s * step)) > 0.5)
    assert int(k.sum()) <= int(pruning(dsc, th=0.5).sum())
    with pytest.raises(ValueError):
        pruning(dsc, delta=0.0)


def test_prune_scores_split_bf16_match_oracle(dscene):
    """Split-BF16 pruning scores (the lattice as a sample stream through the
    renderer's split-BF16 field kernels): within 1e-4 of the FP64 lattice
    evaluation like the FP32 path, and the same keep decisions wherever the
    score is farther than that from the threshold."""
    sc, dsc = dscene
    s3 = prune_scores(dsc, "bf16x3").cpu().numpy()
    s32 = prune_scores(dsc, "fp32").cpu().numpy()
    assert s3.shape == (sc.num_voxels,) and (s3 >= 0).all()
    scale_all = np.maximum(1.0, np.abs(s32))
    assert (np.abs(s3 - s32) <= 1e-4 * scale_all).all()
    rng = np.random.default_rng(4)
    sel = np.sort(rng.choice(sc.num_voxels, 96, replace=False))
    ref = O.prune_max_sigma(O.shift_centers(sc.points, sc.voxel_size), sc.feats, sc.values,
                            sc.voxel_size, sc.mlp, sel)
    scale = np.maximum(1.0, np.abs(ref))
    assert (np.abs(s3[sel] - ref) <= 1e-4 * scale).all()
    clear = np.abs(ref - LN2) > 1e-4 * scale
    keep_ref = (1 - np.exp(-ref)) > 0.5
    keep = (1 - np.exp(-s3[sel].astype(np.float64))) > 0.5
    assert np.array_equal(keep[clear], keep_ref[clear])
    # every lattice point is evaluated on its own: other chunkings (a
    # partial last chunk, one chunk per voxel, one chunk for all) give the
    # same bits
    for cv in (256, 1, sc.num_voxels):
        assert np.array_equal(prune_scores(dsc, "bf16x3", x3_chunk=cv).cpu().numpy(), s3), cv
