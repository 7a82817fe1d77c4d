"""Train-set voxel statistics (encoder.py:594-603 track_voxel_probs, called at
renderer.py:250-251 after each pass) and the pruning that uses them
(encoder.py:612-615): the running per-voxel max of each ray's summed sample
probabilities, against the oracle over the same kept streams, bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell, shell_rays
from nsvf_amd._lib import lib
from nsvf_amd.render import DeviceScene, Renderer
from nsvf_amd.voxels import pruning

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def padded(cnt, off, idx, p):
    W = max(int(cnt.max()), 1)
    I = np.full((len(cnt), W), -1, np.int64)
    Pm = np.zeros((len(cnt), W), np.float32)
    for r in range(len(cnt)):
        I[r, :cnt[r]] = idx[off[r]:off[r] + cnt[r]]
        Pm[r, :cnt[r]] = p[off[r]:off[r] + cnt[r]]
    return I, Pm


@pytest.mark.parametrize("precision,compat", [("fp32", True), ("bf16", False)])
def test_tracked_probs_match_oracle(device, precision, compat):
    sc = shell()
    dsc = DeviceScene(sc, device)
    r = Renderer(device, precision=precision, hierarchical=True, compat=compat)
    r.track = torch.zeros(sc.num_voxels, dtype=torch.float32, device=device)
    ref = np.zeros(sc.num_voxels, np.float32)
    for seed in (7, 8):
        rs, rd = shell_rays(64, seed=seed)
        r.render(dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device), keep=True)
        k = r.kept
        for cnt, off, smp, prob in ((k["c_cnt"], k["c_off"], k["c_samples"], k["c_prob"]),
                                    (k["f_cnt"], k["f_off"], k["f_samples"], k["f_prob"])):
            I, Pm = padded(N(cnt), N(off), N(smp[0]), N(prob))
            ref = O.track_voxel_probs(I, Pm, sc.num_voxels, ref)
    got = N(r.track)
    assert (got > 0).sum() > 100
    assert np.array_equal(got, ref)
    keep = pruning(dsc, th=0.5, train_stats=True, max_voxel_probs=r.track)
    assert np.array_equal(N(keep), ref > 0.5)


def test_tracking_repeated_ids_and_long_rows(device):
    """A voxel id that comes back later in a ray (summed over both runs), -1
    samples, and a ray with more runs than the kernel stages (its lane-0
    path), through the C entry on a hand-made stream."""
    rng = np.random.default_rng(0)
    rows = [[3, 3, 5, -1, 3, 7], [5, 5, 5], list(np.tile([1, 2], 150)), [-1, -1]]
    cnt = np.array([len(x) for x in rows], np.int32)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    idx = np.concatenate([np.asarray(x, np.int32) for x in rows])
    p = rng.random(len(idx)).astype(np.float32)
    rec = np.zeros((len(idx), 4), np.int32)
    rec[:, 0] = idx
    nv = 9
    prev = rng.random(nv).astype(np.float32) * 0.2
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    hl, nh = dev(np.arange(len(rows), dtype=np.int32)), dev(np.array([len(rows)], np.int32))
    vmax = dev(prev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    assert lib().nsvf_track_voxel_probs(P(hl), P(nh), len(rows), P(dev(off)), P(dev(cnt)),
                                        P(dev(rec)), P(dev(p)), P(vmax), nv, st) == 0
    I, Pm = padded(cnt, off, idx, p)
    assert np.array_equal(N(vmax), O.track_voxel_probs(I, Pm, nv, prev))
