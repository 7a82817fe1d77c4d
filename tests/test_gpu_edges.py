"""Edge cases of the fused renderer against the oracle: a view that hits
nothing, a one-voxel scene (the octree's root is the leaf), voxels far
apart (chains of one-child nodes down to each), rays that start inside
voxels (entry depth 0), ray counts that are not a multiple of a wave, and a
one-hit row budget."""
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


def test_far_apart_voxels(device):
    """Three voxels far apart: each hangs below the root on a chain of
    one-child nodes, which the packed records skip, so the root's record
    holds the leaves themselves; hits, their order along each ray and the
    image against the oracle."""
    pts = np.array([[-3.0, -3.0, -3.0], [0.0, 0.25, 0.0], [3.0, 3.0, 2.75]], np.float32)
    sc = IO.scene_from_points(pts, 0.25, embed_dim=32)
    rng = np.random.default_rng(4)
    rs = np.tile(np.array([[-6.0, -6.2, -5.9]], np.float32), (90, 1))
    tgt = pts[rng.integers(0, 3, 90)] + rng.uniform(-0.1, 0.1, (90, 3)).astype(np.float32)
    rd = (tgt - rs).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    root = DeviceScene(sc, device).tree_records[-2].cpu().numpy()   # node T - 1's record
    ids, scale = root[1::2, 3].view(np.int32), root[0::2, 3].view(np.int32)
    assert (ids >= 0).sum() == 3 and (scale[ids >= 0] == 1).all()    # the three leaves
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


def sampler_pair(r, device):
    """The coarse weights and fine stream of the view r rendered last (kept),
    and the same from the two-kernel sequence the fine sampler folds in:
    nsvf_render_composite's weights, then nsvf_render_sample_fine without
    c_field reading them."""
    from nsvf_amd._lib import check, lib
    from nsvf_amd.render import _p, _stream
    L, ws, k = lib(), r._ws, r.kept
    R = len(k["count"])
    n_hit = len(k["hit_list"])
    scal, st = ws["scalars"], _stream(device)
    prob = torch.full_like(k["c_prob"], float("nan"))
    tmp = torch.empty(R * 5, dtype=torch.float32, device=device)
    check(L.nsvf_render_composite(_p(ws["hit_list"]), _p(scal[0:1]), n_hit, _p(ws["c_off"]),
                                  _p(ws["c_cnt"]), _p(ws["c_samples"]), _p(ws["c_field"]),
                                  _p(prob), _p(tmp[: R * 3]), _p(tmp[R * 3: R * 4]),
                                  _p(tmp[R * 4:]), st), "composite")
    fs = torch.full_like(ws["f_samples"], -7)
    f_cnt = torch.full_like(ws["f_cnt"], -7)
    src = torch.full_like(ws["f_src"], -7)
    check(L.nsvf_render_sample_fine(_p(ws["hit_list"]), _p(scal[0:1]), R, _p(ws["c_off"]),
                                    _p(ws["c_cnt"]), _p(scal[2:3]), _p(ws["c_samples"]), None,
                                    _p(prob), _p(ws["f_bound"]), _p(ws["f_off"]), r.fine_steps,
                                    r.compat, _p(fs), _p(f_cnt), _p(src), None, 0, None, -1,
                                    st), "sample_fine")
    torch.cuda.synchronize()
    hl = N(k["hit_list"]).astype(np.int64)
    m = len(k["f_samples"][0])
    return (N(prob), N(fs)[: m * 4].reshape(m, 4), N(f_cnt)[hl], N(src), N(ws["f_bound"])[hl],
            N(ws["f_src"]))


@pytest.mark.parametrize("scene", ["shell", "cube"])
@pytest.mark.parametrize("compat", [False, True])
def test_fine_sampler_composites_the_coarse_rows(device, cube, scene, compat):
    """The fine sampler composites each ray's coarse row itself (no coarse
    composite runs in a hierarchical view): its weights, fine stream, counts
    and reuse marks equal the two-kernel sequence's bit for bit -- on the
    shell and on the cube's rows of more than 128 coarse samples (weights
    read back) and of more than 384 (the sequential walk)."""
    if scene == "shell":
        sc = shell()
        rng = np.random.default_rng(6)
        rs = np.tile(np.array([[0.2, -2.5, 1.1]], np.float32), (3000, 1))
        tgt = rng.uniform(-0.8, 0.8, (3000, 3)).astype(np.float32)
        rd = (tgt - rs).astype(np.float32)
        rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    else:
        sc, rs, rd = cube
    dsc = DeviceScene(sc, device)
    rs_d = torch.from_numpy(np.ascontiguousarray(rs)).to(device)
    rd_d = torch.from_numpy(np.ascontiguousarray(rd)).to(device)
    unfolded = Renderer(device, precision="bf16", hierarchical=True, compat=compat,
                        fold_composite=False)
    want = [N(t) for t in unfolded.render(dsc, rs_d, rd_d)]
    r = Renderer(device, precision="bf16", hierarchical=True, compat=compat)
    out = [N(t) for t in r.render(dsc, rs_d, rd_d, keep=True)]
    for a, b in zip(out, want):  # the image, bit for bit
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    k = r.kept
    prob, fs, f_cnt, src, f_bound, src_got = sampler_pair(r, device)
    c_off, c_cnt = N(k["c_off"]), N(k["c_cnt"])
    if scene == "cube":
        assert (c_cnt > 384).sum() > 10
    got_p = N(k["c_prob"])
    for o, n in zip(c_off, c_cnt):
        assert np.array_equal(got_p[o: o + n].view(np.int32), prob[o: o + n].view(np.int32))
    f_off = N(k["f_off"])
    assert np.array_equal(N(k["f_cnt"]), f_cnt)
    got = N(k["f_samples"][0]), N(k["f_samples"][1]), N(k["f_samples"][2])
    for o, n, b in zip(f_off, f_cnt, f_bound):
        assert np.array_equal(got[0][o: o + n], fs[o: o + n, 0])
        assert np.array_equal(got[1][o: o + n].view(np.int32), fs[o: o + n, 1])
        assert np.array_equal(got[2][o: o + n].view(np.int32), fs[o: o + n, 2])
        assert np.array_equal(src_got[o: o + b], src[o: o + b])
