"""Training-mode sampling noise on the GPU against the oracle: both fused
samplers, with a noise seed, write the same sample streams as the oracle's
sampler fed with the same counter-based noise (hash_noise rows of the hit
rays' frame ids), bit for bit, in compat and strict mode, for the coarse
pass and for the fine pass given the GPU's coarse weights.  A seed of -1
(none) leaves the evaluation sampler unchanged; the training renderer draws
a new stream every step unless deterministic_step is set."""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell, shell_rays
from nsvf_amd.render import DeviceScene, Renderer
from nsvf_amd.train import TrainableField, TrainRenderer

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def csr_to_padded(cnt, off, arr, width, fill):
    out = np.full((len(cnt), width), fill, dtype=arr.dtype)
    for n in range(len(cnt)):
        out[n, : cnt[n]] = arr[off[n]: off[n] + cnt[n]]
    return out


@pytest.fixture(scope="module")
def setup(device):
    sc = shell()
    rs, rd = shell_rays(width=48, seed=11)
    return sc, rs, rd, DeviceScene(sc, device)


def render_kept(dsc, rs, rd, device, seed, compat):
    r = Renderer(device, precision="fp32", hierarchical=True, compat=compat, noise_seed=seed)
    print(f"noise seed {seed}, compat {compat}, rays {rs.shape[0]}", flush=True)
    r.render(dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device), keep=True)
    torch.cuda.synchronize()
    return r.kept


def stream(k, prefix, width):
    cnt, off = N(k[f"{prefix}_cnt"]), N(k[f"{prefix}_off"])
    s = k[f"{prefix}_samples"]
    return (csr_to_padded(cnt, off, N(s[0]), width, -1),
            csr_to_padded(cnt, off, N(s[1]), width, np.float32(O.MAX_DEPTH)),
            csr_to_padded(cnt, off, N(s[2]), width, np.float32(0)), cnt)


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("seed", [0, 12345, 2**32 - 1])
def test_noisy_streams_bit_exact(setup, device, seed, compat):
    sc, rs, rd, dsc = setup
    k = render_kept(dsc, rs, rd, device, seed, compat)
    # the coarse stream depends only on geometry and noise
    ref = O.render_view(sc, rs, rd, compat=compat, noise_seed=seed, hierarchical=False)
    ci = ref["c_idx"]
    si, sd, sx, cnt = stream(k, "c", ci.shape[1])
    assert np.array_equal(cnt, (ci != -1).sum(1))
    assert np.array_equal(si, ci)
    assert np.array_equal(sd, ref["c_depth"])
    assert np.array_equal(sx, ref["c_dist"])
    # the fine stream given the GPU's coarse weights
    c_cnt, c_off = N(k["c_cnt"]), N(k["c_off"])
    gp = csr_to_padded(c_cnt, c_off, N(k["c_prob"])[: c_off[-1] + c_cnt[-1]], ci.shape[1],
                       np.float32(0))
    ref2 = O.render_view(sc, rs, rd, compat=compat, noise_seed=seed, coarse_probs=gp)
    fi = ref2["f_idx"]
    fsi, fsd, fsx, f_cnt = stream(k, "f", fi.shape[1])
    assert np.array_equal(f_cnt, (fi != -1).sum(1))
    assert np.array_equal(fsi, fi)
    assert np.array_equal(fsd, ref2["f_depth"])
    assert np.array_equal(fsx, ref2["f_dist"])


def test_seed_changes_samples_and_none_is_deterministic(setup, device):
    sc, rs, rd, dsc = setup
    a = render_kept(dsc, rs, rd, device, 5, False)
    a = [N(t).copy() for t in a["c_samples"][:3]]
    b = render_kept(dsc, rs, rd, device, 5, False)
    assert all(np.array_equal(x, N(y)) for x, y in zip(a, b["c_samples"][:3]))
    c = render_kept(dsc, rs, rd, device, 6, False)
    assert not np.array_equal(a[1], N(c["c_samples"][1]))
    d = render_kept(dsc, rs, rd, device, None, False)
    ref = O.render_view(sc, rs, rd, compat=False, hierarchical=False)
    ci = ref["c_idx"]
    si, sd, _, _ = stream(d, "c", ci.shape[1])
    assert np.array_equal(si, ci) and np.array_equal(sd, ref["c_depth"])


def test_train_renderer_draws_new_noise_each_step(setup, device):
    sc, rs, rd, dsc = setup
    t = lambda a: torch.from_numpy(a).to(device)
    field = TrainableField(sc, device)
    tr = TrainRenderer(device, seed=40)
    with torch.no_grad():
        tr.forward(dsc, field, t(rs), t(rd))
        s1 = N(tr.last.samples).copy()
        assert tr.last_seed == 40
        tr.forward(dsc, field, t(rs), t(rd))
        assert tr.last_seed == 41
        assert not np.array_equal(s1, N(tr.last.samples))
        tr.forward(dsc, field, t(rs), t(rd), seed=40)
        assert np.array_equal(s1, N(tr.last.samples))
        det = TrainRenderer(device, deterministic_step=True)
        det.forward(dsc, field, t(rs), t(rd))
        assert det.last_seed is None
        ev = Renderer(device, precision="fp32", hierarchical=False)
        g = ev.coarse_stream(dsc, t(rs), t(rd))
        assert torch.equal(det.last.samples, g.samples[: det.last.samples.shape[0]])
