"""Differentiation on the GPU: the gather and compositing backward kernels
against the oracle's FP64 gradients, and the training loss's gradients
(4 views x 2048 rays of the spherical-shell scene, MSE against uniform
targets) against the plain path's torch autograd and against FP64 central
differences of the oracle's loss.

Tolerances: kernel gradients within 1e-5 of the FP64 oracle (relative to the
largest entry); the training gradients within 1e-4 (relative Frobenius
norm) of the plain path's FP32 autograd, and within 1e-4 relative of the
central differences (eps 1e-4; eps 1e-6 for parameters whose 1e-4 interval
straddles a ReLU kink), with an absolute floor of 1e-6 times the largest
difference for entries whose gradient nearly cancels."""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell, shell_rays
from nsvf_amd import scene as S
from nsvf_amd.render import DeviceScene, Renderer
from nsvf_amd.train import (StreamComposite, TrainableField, TrainRenderer, gather_backward_plan,
                            mse_loss, trilinear_gather, trilinear_gather_backward)

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def stream_to_padded(g):
    """Padded [n_hit, S] (idx, depth, dist) of a coarse stream, hit rays in
    frame order, plus the slot index of each padded entry (-1 at pads)."""
    hl = N(g.hit_list[:g.n_hit]).astype(np.int64)
    off = N(g.off)[hl].astype(np.int64)
    cnt = N(g.cnt)[hl].astype(np.int64)
    rec = N(g.samples)
    W = int(cnt.max()) if len(cnt) else 0
    slot = np.full((len(hl), W), -1, np.int64)
    for n in range(len(hl)):
        slot[n, :cnt[n]] = off[n] + np.arange(cnt[n])
    ok = slot >= 0
    si = np.where(ok, rec[np.maximum(slot, 0), 0], -1).astype(np.int32)
    sd = np.where(ok, rec[np.maximum(slot, 0), 1].view(np.float32), O.MAX_DEPTH).astype(np.float32)
    sx = np.where(ok, rec[np.maximum(slot, 0), 2].view(np.float32), 0).astype(np.float32)
    return hl, si, sd, sx, slot


@pytest.fixture(scope="module")
def scene(device):
    sc = shell()
    return sc, DeviceScene(sc, device)


# ---------------------------------------------------------------------------
# kernels

def test_gather_backward_matches_oracle_and_is_deterministic(scene, device):
    sc, dsc = scene
    rng = np.random.default_rng(21)
    M, D = 30000, sc.values.shape[1]
    cen = O.shift_centers(sc.points, sc.voxel_size)
    vi = rng.integers(0, sc.num_voxels, M).astype(np.int32)
    vi[rng.random(M) < 0.1] = -1
    # many samples in a few voxels too (long segments)
    vi[:3000] = rng.integers(0, 5, 3000)
    x = (cen[np.maximum(vi, 0)] + rng.uniform(-0.5, 0.5, (M, 3)).astype(np.float32)
         * np.float32(sc.voxel_size)).astype(np.float32)
    G = rng.standard_normal((M, D)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    args = (t(vi), t(x), dsc.centers, dsc.feats, sc.num_keys, t(G), sc.voxel_size)
    g1 = trilinear_gather_backward(*args)
    g2 = trilinear_gather_backward(*args)
    # with the scene's precomputed incidence order: the same bits
    g3 = trilinear_gather_backward(*args, gather_backward_plan(dsc.feats, sc.num_keys))
    torch.cuda.synchronize()
    assert torch.equal(g1, g2), "backward is not deterministic"
    assert torch.equal(g1, g3), "the planned backward differs"
    ref = O.trilinear_backward(vi, x, cen, sc.feats, sc.num_keys, G, sc.voxel_size)
    err = np.abs(N(g1) - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err
    untouched = np.ones(sc.num_keys, bool)
    untouched[sc.feats[vi[vi >= 0]].ravel()] = False
    assert (N(g1)[untouched] == 0).all()


def test_gather_forward_zero_rows_for_unused_slots(scene, device):
    sc, dsc = scene
    vi = torch.tensor([-1, 5, -1, 17], dtype=torch.int32, device=device)
    cen = O.shift_centers(sc.points, sc.voxel_size)
    x = torch.from_numpy(cen[[0, 5, 0, 17]].copy()).to(device)
    vals = torch.from_numpy(sc.values).to(device)
    out = N(trilinear_gather(vals, vi, x, dsc.centers, dsc.feats, sc.voxel_size))
    assert (out[[0, 2]] == 0).all()
    ref = O.trilinear_gather([5, 17], cen[[5, 17]], cen, sc.feats, sc.values, sc.voxel_size,
                             dtype=np.float32)
    assert np.array_equal(out[[1, 3]], ref)


def test_gather_backward_empty_and_bad_arguments(scene, device):
    sc, dsc = scene
    D = sc.values.shape[1]
    e = torch.zeros(0, dtype=torch.int32, device=device)
    g = trilinear_gather_backward(e, torch.zeros((0, 3), device=device), dsc.centers, dsc.feats,
                                  sc.num_keys, torch.zeros((0, D), device=device), sc.voxel_size)
    assert g.shape == (sc.num_keys, D) and not g.any()
    with pytest.raises(RuntimeError, match="multiple of 4"):
        trilinear_gather_backward(e, torch.zeros((0, 3), device=device), dsc.centers, dsc.feats,
                                  sc.num_keys, torch.zeros((0, 6), device=device), sc.voxel_size)


def test_composite_backward_matches_oracle(scene, device):
    sc, dsc = scene
    rs, rd = shell_rays(width=48, seed=5)
    R = rs.shape[0]
    rnd = Renderer(device, precision="fp32", hierarchical=False)
    g = rnd.coarse_stream(dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device))
    g = g.owned()
    rng = np.random.default_rng(8)
    field = np.zeros((g.slots, 4), np.float32)
    field[:, 0] = rng.normal(2.0, 6.0, g.slots)
    field[:, 1:] = rng.uniform(0, 1, (g.slots, 3))
    gr = rng.standard_normal((R, 3)).astype(np.float32)
    gd = rng.standard_normal(R).astype(np.float32)
    gm = rng.standard_normal(R).astype(np.float32)
    f = torch.from_numpy(field).to(device).requires_grad_(True)
    rgb, dep, mis = StreamComposite.apply(f, g)
    loss = ((rgb * torch.from_numpy(gr).to(device)).sum() + (dep * torch.from_numpy(gd).to(device)).sum()
            + (mis * torch.from_numpy(gm).to(device)).sum())
    loss.backward()
    gf = N(f.grad)
    hl, si, sd, sx, slot = stream_to_padded(g)
    ok = slot >= 0
    sig = np.where(ok, field[np.maximum(slot, 0), 0], 0)
    col = np.where(ok[..., None], field[np.maximum(slot, 0), 1:], 0)
    ref_s, ref_c = O.composite_backward(sig, col, sd, sx, si, gr[hl], gd[hl], gm[hl])
    got_s = np.where(ok, gf[np.maximum(slot, 0), 0], 0)
    got_c = np.where(ok[..., None], gf[np.maximum(slot, 0), 1:], 0)
    assert np.abs(got_s - ref_s).max() < 1e-5 * max(1.0, np.abs(ref_s).max())
    assert np.abs(got_c - ref_c).max() < 1e-5 * max(1.0, np.abs(ref_c).max())
    # forward outputs against the oracle too
    c_ref, d_ref, m_ref, _ = O.composite(sig, col, sd, sx, si)
    assert np.abs(N(rgb)[hl] - c_ref).max() < 1e-5
    assert np.abs(N(mis)[hl] - m_ref).max() < 1e-5
    # unused slots get no gradient
    used = np.zeros(g.slots, bool)
    used[slot[ok]] = True
    assert (gf[~used] == 0).all()


# ---------------------------------------------------------------------------
# the training step

def c3_batch():
    """4 views x 2048 pixels of 64x64 (seed 2), targets U(0,1) (seed 3)."""
    rng = np.random.default_rng(2)
    rs_all, rd_all = [], []
    for view in range(4):
        cam = S.sphere_camera(2.5, 100 + view)
        rs, rd = S.camera_rays(64, 64, cam)
        sel = rng.choice(rs.shape[0], 2048, replace=False)
        rs_all.append(rs[sel])
        rd_all.append(rd[sel])
    rs = np.concatenate(rs_all).astype(np.float32)
    rd = np.concatenate(rd_all).astype(np.float32)
    target = np.random.default_rng(3).uniform(0, 1, (rs.shape[0], 3)).astype(np.float32)
    return rs, rd, target


@pytest.fixture(scope="module")
def c3(scene, device):
    sc, dsc = scene
    rs, rd, target = c3_batch()
    t = lambda a: torch.from_numpy(a).to(device)
    field = TrainableField(sc, device)
    tr = TrainRenderer(device, deterministic_step=True)
    rgb, _, _ = tr.forward(dsc, field, t(rs), t(rd))
    loss = mse_loss(rgb, t(target))
    loss.backward()
    torch.cuda.synchronize()
    grads = {"values": N(field.values.grad)}
    for i, (w, b) in enumerate(zip(field.weights, field.biases)):
        grads[f"w{i}"] = N(w.grad)
        grads[f"b{i}"] = N(b.grad)
    return {"rs": rs, "rd": rd, "target": target, "loss": float(loss.detach()), "grads": grads,
            "rgb": N(rgb), "n_hit": tr.last.n_hit}


def test_training_gradients_match_plain_autograd(scene, c3, device):
    from nsvf_amd.plain import PlainNSVF
    sc, _ = scene
    pl = PlainNSVF(sc, device, hierarchical=False)
    pl.values.requires_grad_(True)
    for w, b in pl.layers:
        w.requires_grad_(True)
        b.requires_grad_(True)
    t = lambda a: torch.from_numpy(a).to(device)
    rgb, _, _ = pl.forward(t(c3["rs"]), t(c3["rd"]))
    loss = torch.mean((rgb - t(c3["target"])) ** 2)
    loss.backward()
    lv = loss.item()
    assert abs(lv - c3["loss"]) <= 1e-6 * abs(lv)
    assert np.abs(N(rgb) - c3["rgb"]).max() < 1e-5
    ref = {"values": N(pl.values.grad)}
    for i, (w, b) in enumerate(pl.layers):
        ref[f"w{i}"] = N(w.grad)
        ref[f"b{i}"] = N(b.grad)
    for k, g in c3["grads"].items():
        rel = np.linalg.norm(g - ref[k]) / max(np.linalg.norm(ref[k]), 1e-30)
        assert rel < 1e-4, (k, rel)


def step_grads(scene, c3, device, **kw):
    sc, dsc = scene
    t = lambda a: torch.from_numpy(a).to(device)
    field = TrainableField(sc, device, mlp=kw.pop("mlp", "hip"))
    tr = TrainRenderer(device, deterministic_step=True, **kw)
    rgb, _, _ = tr.forward(dsc, field, t(c3["rs"]), t(c3["rd"]))
    loss = mse_loss(rgb, t(c3["target"]))
    loss.backward()
    torch.cuda.synchronize()
    grads = {"values": N(field.values.grad)}
    for i, (w, b) in enumerate(zip(field.weights, field.biases)):
        grads[f"w{i}"] = N(w.grad)
        grads[f"b{i}"] = N(b.grad)
    return float(loss.detach()), N(rgb), grads


def test_batch_order_traversal_gives_the_same_step(scene, c3, device):
    """The traversal's grouping of the batch's rays (coherent, the default,
    against batch order) changes nothing: loss, image and every gradient
    are the same bits."""
    loss, rgb, grads = step_grads(scene, c3, device, coherent=False)
    assert loss == c3["loss"] and np.array_equal(rgb, c3["rgb"])
    for k, g in c3["grads"].items():
        assert np.array_equal(grads[k], g), k


def test_split_bf16_training_step(scene, c3, device):
    """A step with the MLP in split BF16 (mlp="hip_x3": the renderer's
    split-BF16 kernel forward; the backward's recomputation and input
    and parameter gradients in split BF16, the head's recomputation in FP32),
    against the FP32 step: image within 1e-4 and loss within 1e-5
    (relative), MLP gradients within 1e-3 (relative Frobenius norm).
    Where a pre-activation lies within the two forwards' difference (~1e-6)
    of a ReLU kink the steps take different sides and that sample's
    gradient differs by a whole term.  At the compositing's relu(sigma) this
    happens to two of the 525k samples on this batch, each feeding only its
    voxel's 8 corner embeddings, which hold most of the embedding gradient's
    difference (1.2e-3 relative): the test takes out the 16 rows that differ
    most, and holds the rest -- where the hidden layers' kinks still act,
    spread over many samples -- to 1e-3."""
    loss, rgb, grads = step_grads(scene, c3, device, mlp="hip_x3")
    print(f"hip_x3 step: loss {loss:.8f} vs {c3['loss']:.8f}, "
          f"max |d rgb| {np.abs(rgb - c3['rgb']).max():.2e}", flush=True)
    assert abs(loss - c3["loss"]) <= 1e-5 * abs(c3["loss"])
    assert np.abs(rgb - c3["rgb"]).max() < 1e-4
    rels = {k: float(np.linalg.norm(grads[k] - g) / max(np.linalg.norm(g), 1e-30))
            for k, g in c3["grads"].items()}
    g, ref = grads["values"], c3["grads"]["values"]
    keep = np.argsort(np.linalg.norm(g - ref, axis=1))[:-16]
    rels["values_rest"] = float(np.linalg.norm(g[keep] - ref[keep]) / np.linalg.norm(ref[keep]))
    print("hip_x3 gradients, relative to FP32:",
          " ".join(f"{k} {v:.2e}" for k, v in rels.items()), flush=True)
    for k, rel in rels.items():
        if k != "values":
            assert rel < 1e-3, (k, rel)
    assert rels["values"] < 1e-2


@pytest.fixture(scope="module")
def fd_checks(scene, c3):
    """64 parameters of the C3 step with FP64 central differences of the
    oracle's loss on the same samples at eps 1e-4 and 1e-6: (name, position,
    difference at 1e-4, difference at 1e-6)."""
    sc, _ = scene
    rs, rd, target = c3["rs"], c3["rd"], c3["target"]
    ref = O.render_view(sc, rs, rd, hierarchical=False)
    assert int(ref["hits"].sum()) == c3["n_hit"]
    obj = O.FixedSampleLoss(ref["c_idx"], ref["c_depth"], ref["c_dist"], ref["hits"], rs, rd,
                            O.shift_centers(sc.points, sc.voxel_size), sc.feats, sc.voxel_size,
                            target)
    R3 = 3 * rs.shape[0]
    # the objective restates the forward: its loss equals the frame MSE up to
    # the constant of the rays that miss
    miss = ~ref["hits"]
    const = float(((1.0 - target[miss].astype(np.float64)) ** 2).sum())
    full = (obj.loss(sc.values, sc.mlp) + const) / R3
    assert abs(full - c3["loss"]) < 1e-6 * full
    # Central differences at eps 1e-4 and 1e-6.  The loss is piecewise smooth
    # (ReLU in the MLP and on sigma): where the eps 1e-4 interval straddles a
    # kink the two differences disagree, and the parameter is checked at 1e-6
    # instead, where FP64 rounding still leaves ~1e-10 relative accuracy.
    EPS, FINE = 1e-4, 1e-6
    rng = np.random.default_rng(4)
    checks = []
    # 48 embedding entries of keys that samples touch
    touched = np.unique(obj.keys)
    for k, d in zip(rng.choice(touched, 48, replace=False), rng.integers(0, sc.values.shape[1], 48)):
        rows = obj.rows_touching_key(k)

        def fd(eps):
            vp, vm = sc.values.astype(np.float64), sc.values.astype(np.float64)
            vp[k, d] += eps
            vm[k, d] -= eps
            return (obj.loss(vp, sc.mlp, rows) - obj.loss(vm, sc.mlp, rows)) / (2 * eps) / R3
        checks.append(("values", (k, d), fd(EPS), fd(FINE)))
    # 16 MLP entries: 3 weights and 1 bias of each layer
    emb = obj.embed(sc.values)
    n_layers = len(sc.mlp.weights)
    for i in range(n_layers):
        W = sc.mlp.weights[i]
        picks = [("w", (int(rng.integers(W.shape[0])), int(rng.integers(W.shape[1])))) for _ in range(3)]
        picks.append(("b", (int(rng.integers(W.shape[0])),)))
        for kind, pos in picks:
            a, b = (obj.mlp_central_difference(emb, sc.mlp, i, kind, pos, e) / R3 for e in (EPS, FINE))
            checks.append((f"{kind}{i}", pos, a, b))
    assert len(checks) == 64
    return checks


def check_against_fd(checks, g, floor_rel):
    """Every gradient within 1e-4 (relative) + floor_rel of the largest of
    the FP64 difference: the eps 1e-4 one, or where that interval straddles
    a kink (the two differences disagree) the eps 1e-6 one."""
    big = max(abs(c[3]) for c in checks)
    kink = [abs(c[2] - c[3]) > 1e-4 * abs(c[3]) + 1e-6 * big for c in checks]
    kinked = [c for c, k in zip(checks, kink) if k]
    got = [g[c[0]][c[1]] for c in checks]
    diff = [abs(x - (c[3] if k else c[2])) for c, k, x in zip(checks, kink, got)]
    bad = [(c, x) for c, x, d in zip(checks, got, diff) if d > 1e-4 * abs(c[3]) + floor_rel * big]
    print(f"{len(kinked)} of 64 parameters sit within 1e-4 of a ReLU kink; largest difference "
          f"{max(diff) / big:.1e} of the largest gradient", flush=True)
    assert not bad, bad
    assert len(kinked) <= 12, kinked


def test_training_gradients_match_fp64_central_differences(fd_checks, c3):
    check_against_fd(fd_checks, c3["grads"], 1e-6)


def test_split_bf16_training_gradients_match_fp64_central_differences(scene, c3, fd_checks,
                                                                      device):
    """The split-BF16 step's gradients (mlp="hip_x3") at the same 64
    parameters against the same FP64 central differences: within 1e-4
    relative + 5e-5 of the largest of them (the FP32 step: + 1e-6).  Its
    input-gradient products carry ~2^-16 relative error per term, and a
    weight gradient that is a small sum of large terms over 525k samples
    keeps their absolute error (1.8e-5 of the largest gradient measured on
    a W_0 entry 800x smaller than it)."""
    check_against_fd(fd_checks, step_grads(scene, c3, device, mlp="hip_x3")[2], 5e-5)
