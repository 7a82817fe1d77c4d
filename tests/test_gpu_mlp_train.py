"""The training MLP kernels (FP32 MFMA forward and backward) against float64
autograd of the same MLP on the CPU (field.py:217-279: Linear + ReLU
layers, a 4-wide head, output (sigma, sigmoid(rgb))).

Bars: outputs within 2e-6 absolute (the output is O(1)); input and
parameter gradients within 2e-5 of the largest float64 entry of each
tensor (FP32 sums over 128 features and, for the parameters, over all
samples); the backward is bit-for-bit repeatable.  Samples with a hidden
pre-activation within 1e-5 of zero get a zero output gradient: there FP32
and float64 may take different sides of a ReLU kink, where the gradient
jumps (about 2 of 15M units at 38k samples)."""
import numpy as np
import pytest
import torch

from nsvf_amd.field import pack_mlp
from nsvf_amd.scene import FieldMLP as SceneMLP
from nsvf_amd.train import FieldMLP, FieldMLPx3, mlp_train_supported, pack_mlp_train, pack_mlp_x3

pytestmark = pytest.mark.gpu

D, H, L = 32, 128, 3


def params(seed, device):
    rng = np.random.default_rng(seed)
    dims = [D] + [H] * L + [4]
    out = []
    for i in range(L + 1):
        w = rng.normal(0, np.sqrt(2.0 / dims[i]), (dims[i + 1], dims[i])).astype(np.float32)
        b = rng.normal(0, 0.1, dims[i + 1]).astype(np.float32)
        out += [torch.from_numpy(w).to(device), torch.from_numpy(b).to(device)]
    return out


def reference(emb, ps, dout):
    """float64 autograd on the CPU: (out, d emb, d params)."""
    e = emb.detach().cpu().double().requires_grad_(True)
    q = [p.detach().cpu().double().requires_grad_(True) for p in ps]
    h = e
    for i in range(L + 1):
        h = torch.nn.functional.linear(h, q[2 * i], q[2 * i + 1])
        if i < L:
            h = torch.relu(h)
    out = torch.cat([h[:, :1], torch.sigmoid(h[:, 1:4])], 1)
    out.backward(dout.detach().cpu().double())
    return out.detach(), e.grad, [t.grad for t in q]


def off_kinks(emb, ps, dout):
    """dout with the rows of samples near a ReLU kink zeroed."""
    h = emb.detach().cpu().double()
    near = torch.zeros(h.shape[0], dtype=torch.bool)
    for i in range(L):
        z = torch.nn.functional.linear(h, ps[2 * i].cpu().double(), ps[2 * i + 1].cpu().double())
        near |= (z.abs() < 1e-5).any(1)
        h = torch.relu(z)
    out = dout.clone()
    out[near.to(dout.device)] = 0.0
    return out, int(near.sum())


def run(emb, ps, dout, fn=FieldMLP):
    e = emb.clone().requires_grad_(True)
    q = [p.clone().requires_grad_(True) for p in ps]
    out = fn.apply(e, *q)
    out.backward(dout)
    torch.cuda.synchronize()
    return out.detach(), e.grad, [t.grad for t in q]


def close(got, want, tol):
    want = want.double()
    scale = max(float(want.abs().max()), 1e-30)
    return float((got.detach().cpu().double() - want).abs().max()) / scale, tol


@pytest.mark.parametrize("M,seed", [(1, 1), (1000, 2), (128 * 300 + 5, 3)])
def test_forward_backward_match_float64(device, M, seed):
    assert mlp_train_supported(D, H, L)
    ps = params(seed, device)
    g = torch.Generator().manual_seed(seed)
    emb = torch.randn((M, D), generator=g).mul_(0.3).to(device)
    dout, n_kink = off_kinks(emb, ps, torch.randn((M, 4), generator=g).to(device))
    print(f"M={M} seed={seed}: {n_kink} samples near a kink", flush=True)
    out, demb, dps = run(emb, ps, dout)
    r_out, r_demb, r_dps = reference(emb, ps, dout)
    assert float((out.cpu().double() - r_out).abs().max()) < 2e-6
    err, tol = close(demb, r_demb, 2e-5)
    assert err < tol, ("demb", err)
    names = [f"{k}{i}" for i in range(L + 1) for k in ("W", "b")]
    for name, a, b in zip(names, dps, r_dps):
        err, tol = close(a, b, 2e-5)
        assert err < tol, (name, err)


def test_backward_is_deterministic_and_empty_input(device):
    ps = params(4, device)
    g = torch.Generator().manual_seed(4)
    emb = torch.randn((5000, D), generator=g).mul_(0.3).to(device)
    dout = torch.randn((5000, 4), generator=g).to(device)
    a = run(emb, ps, dout)
    b = run(emb, ps, dout)
    assert torch.equal(a[1], b[1])
    assert all(torch.equal(x, y) for x, y in zip(a[2], b[2]))
    out, demb, dps = run(emb[:0], ps, dout[:0])
    assert out.shape == (0, 4) and demb.shape == (0, D)
    assert all(not t.any() for t in dps)


def test_pack_rejects_unbuilt_shape(device):
    ps = params(5, device)
    w0 = torch.zeros((H, 48), device=device)
    with pytest.raises(RuntimeError, match="not built"):
        pack_mlp_train([w0] + ps[2::2], ps[1::2])


def test_split_bf16_weight_pack_on_the_device(device):
    """pack_mlp_x3 (torch ops on the device, every training step) equals
    the scene packer's split-BF16 fragments and biases bit for bit."""
    ps = params(6, device)
    frags, bias = pack_mlp_x3(ps[0::2], ps[1::2])
    host = pack_mlp(SceneMLP([p.cpu().numpy() for p in ps[0::2]],
                             [p.cpu().numpy() for p in ps[1::2]]))
    assert torch.equal(frags.cpu(), host["bf16x3"]) and torch.equal(bias.cpu(), host["bias"])


@pytest.mark.parametrize("M,seed", [(1, 1), (1000, 2), (128 * 300 + 5, 3)])
def test_split_bf16_forward_backward(device, M, seed):
    """FieldMLPx3: the features written as the split-BF16 kernel's operand
    tiles (M not a multiple of the tile), its forward within 1e-4 of float64
    (three BF16 products per FP32 one; plain BF16 is ~1e-2 here); its
    backward (recomputation, input and parameter gradients in split BF16;
    the head's recomputed logits and input gradient in FP32) within 1e-4 of
    the largest float64 entry of each gradient off the ReLU kinks (FP32:
    2e-5), and bit-for-bit repeatable."""
    ps = params(seed, device)
    g = torch.Generator().manual_seed(seed)
    emb = torch.randn((M, D), generator=g).mul_(0.3).to(device)
    dout, n_kink = off_kinks(emb, ps, torch.randn((M, 4), generator=g).to(device))
    out, demb, dps = run(emb, ps, dout, FieldMLPx3)
    r_out, r_demb, r_dps = reference(emb, ps, dout)
    err = float((out.cpu().double() - r_out).abs().max())
    errs = {"demb": close(demb, r_demb, 1e-4)[0]}
    names = [f"{k}{i}" for i in range(L + 1) for k in ("W", "b")]
    errs.update({name: close(a, b, 1e-4)[0] for name, a, b in zip(names, dps, r_dps)})
    print(f"M={M}: split-BF16 forward max abs error {err:.2e}; gradients "
          + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert err < 1e-4
    for name, e in errs.items():
        assert e < 1e-4, (name, e)
    _, demb2, dps2 = run(emb, ps, dout, FieldMLPx3)
    assert torch.equal(demb, demb2)
    assert all(torch.equal(a, b) for a, b in zip(dps, dps2))
