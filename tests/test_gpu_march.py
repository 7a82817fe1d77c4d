"""Early ray termination (renderer.py:136-174, --raymarching-tolerance) on
the GPU against the oracle's chunked marching (march_chunks), on the
spherical-shell scene.

Bars: the evaluated samples and the stopped rays equal the oracle's on every
ray whose running total never came within 1e-4 (relative) of the threshold
at a chunk end (there FP32 and FP64 field values may decide differently;
such rays are counted); the FP32 image within 1e-5 of the FP64 oracle on
rays whose evaluated samples agree; with one chunk holding every sample the
image is bit-identical to rendering without termination."""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell, shell_rays
from nsvf_amd.render import DeviceScene, Renderer

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def padded(cnt, off, arr, width, fill):
    out = np.full((len(cnt), width), fill, dtype=arr.dtype)
    for n in range(len(cnt)):
        out[n, : cnt[n]] = arr[off[n]: off[n] + cnt[n]]
    return out


@pytest.fixture(scope="module")
def setup(device):
    sc = shell()
    rs, rd = shell_rays(width=48, seed=9)
    return sc, rs, rd, DeviceScene(sc, device)


def render(dsc, rs, rd, device, hierarchical, tolerance, chunk, precision="fp32"):
    r = Renderer(device, precision=precision, hierarchical=hierarchical, compat=False,
                 tolerance=tolerance, chunk_size=chunk)
    print(f"tolerance {tolerance}, chunk {chunk}, hierarchical {hierarchical}", flush=True)
    out = r.render(dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device),
                   keep=True)
    torch.cuda.synchronize()
    return r, [N(t) for t in out]


def check_pass(k, ref, tag, width):
    cnt, off = N(k[f"{tag}_cnt"]), N(k[f"{tag}_off"])
    ev = padded(cnt, off, N(k[f"{tag}_eval"]).astype(bool), width, False)
    stop = N(k[f"{tag}_active"]) == 0
    want, want_stop, margin = ref[f"{tag}_evaluated"], ref[f"{tag}_stopped"], ref[f"{tag}_margin"]
    sure = margin > 1e-4
    same = (ev == want).all(1)
    assert (same | ~sure).all(), f"{tag}: {(~same & sure).sum()} rays evaluate other samples"
    assert np.array_equal(stop[sure], want_stop[sure])
    print(f"{tag}: {want_stop.sum()} of {len(stop)} rays stop, {(~sure).sum()} marginal, "
          f"{int(want.sum())} samples evaluated", flush=True)
    return same


@pytest.mark.parametrize("chunk", [512, 4096])
def test_coarse_marching_matches_oracle(setup, device, chunk):
    sc, rs, rd, dsc = setup
    r, (rgb, depth, missed) = render(dsc, rs, rd, device, False, 1e-2, chunk)
    ref = O.render_view(sc, rs, rd, hierarchical=False, compat=False, tolerance=1e-2,
                        chunk_size=chunk)
    assert ref["c_stopped"].sum() > 0
    same = check_pass(r.kept, ref, "c", ref["c_idx"].shape[1])
    hit = np.nonzero(ref["hits"])[0][same]
    assert np.abs(rgb[hit] - ref["rgb"][hit]).max() < 1e-5
    assert np.abs(depth[hit] - ref["depth"][hit]).max() < 1e-5
    assert np.abs(missed[hit] - ref["missed"][hit]).max() < 1e-5


def test_hierarchical_marching_matches_oracle(setup, device):
    sc, rs, rd, dsc = setup
    r, (rgb, depth, missed) = render(dsc, rs, rd, device, True, 1e-2, 1024)
    k = r.kept
    c_cnt, c_off = N(k["c_cnt"]), N(k["c_off"])
    W = int(c_cnt.max())
    gp = padded(c_cnt, c_off, N(k["c_prob"]), W, np.float32(0))
    ref = O.render_view(sc, rs, rd, compat=False, tolerance=1e-2, chunk_size=1024,
                        coarse_probs=gp)
    same_c = check_pass(k, ref, "c", ref["c_idx"].shape[1])
    fi = ref["f_idx"]
    f_cnt, f_off = N(k["f_cnt"]), N(k["f_off"])
    assert np.array_equal(padded(f_cnt, f_off, N(k["f_samples"][0]), fi.shape[1], -1), fi)
    same_f = check_pass(k, ref, "f", fi.shape[1])
    hit = np.nonzero(ref["hits"])[0][same_c & same_f]
    assert np.abs(rgb[hit] - ref["rgb"][hit]).max() < 1e-5
    assert np.abs(depth[hit] - ref["depth"][hit]).max() < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_one_chunk_is_no_termination(setup, device, precision):
    sc, rs, rd, dsc = setup
    _, a = render(dsc, rs, rd, device, True, 1e-3, 1 << 30, precision)
    r = Renderer(device, precision=precision, hierarchical=True, compat=False)
    b = [N(t) for t in r.render(dsc, torch.from_numpy(rs).to(device),
                                torch.from_numpy(rd).to(device))]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
