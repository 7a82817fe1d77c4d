"""The compacted-stream entry points on their own: the fine stream's
evaluation list (nsvf_render_eval_list_count + scan + nsvf_render_eval_list)
against numpy on random marks, and the field over a list of stream slots
(nsvf_field_eval / the split kernels with a list) against the field of the
whole stream -- listed slots bit for bit, the others untouched."""
import numpy as np
import pytest
import torch

from _cases import shell, shell_rays
from nsvf_amd._lib import check, lib
from nsvf_amd.render import DeviceScene, Renderer, _p, _stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 63, 4096, 4097, 100003])
def test_eval_list_matches_numpy(device, n):
    rng = np.random.default_rng(n)
    slack = 5000
    marks = rng.choice(np.array([-2, -1, -1, 0, 5, 77, 300], np.int16), n + slack)
    src = torch.from_numpy(marks).to(device)
    n_dev = torch.tensor([n], dtype=torch.int32, device=device)
    n_max = n + slack
    L, st = lib(), _stream(torch.device(device))
    nb = int(L.nsvf_render_eval_list_blocks(n_max))
    bcnt = torch.full((nb,), -7, dtype=torch.int32, device=device)
    boff = torch.empty(nb, dtype=torch.int32, device=device)
    total = torch.zeros(1, dtype=torch.int32, device=device)
    ws = torch.empty(int(L.nsvf_scan_workspace_ints(nb)), dtype=torch.int32, device=device)
    lst = torch.full((n_max,), -9, dtype=torch.int32, device=device)
    check(L.nsvf_render_eval_list_count(_p(src), _p(n_dev), n_max, _p(bcnt), st), "count")
    check(L.nsvf_scan_exclusive_i32(_p(bcnt), _p(boff), nb, _p(ws), _p(total), st), "scan")
    check(L.nsvf_render_eval_list(_p(src), _p(n_dev), n_max, _p(boff), _p(lst), st), "list")
    torch.cuda.synchronize()
    want = np.nonzero(marks[:n] == -1)[0].astype(np.int32)
    E = int(total)
    assert E == len(want)
    got = lst.cpu().numpy()
    assert np.array_equal(got[:E], want)
    assert (got[E:] == -9).all()


@pytest.mark.parametrize("precision,split", [("fp32", None), ("bf16", True), ("bf16", False),
                                             ("bf16x3", None)])
def test_field_over_a_list(device, precision, split):
    sc = shell()
    rs, rd = shell_rays(width=64)
    dsc = DeviceScene(sc, device)
    r = Renderer(device, precision=precision, split_field=split)
    r.render(dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device), keep=True)
    torch.cuda.synchronize()
    m = int(r._ws["scalars"][1])
    samples = r._ws["c_samples"][: m * 4].view(m, 4).clone()
    full = r.kept["c_field"][:m].clone()
    rs_d, rd_d = torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device)
    rng = np.random.default_rng(7)
    sel = np.sort(rng.choice(m, m // 3, replace=False)).astype(np.int32)
    lst = torch.from_numpy(sel).to(device)
    n_dev = torch.tensor([len(sel)], dtype=torch.int32, device=device)
    out = torch.full((m, 4), float("nan"), dtype=torch.float32, device=device)
    r._field(dsc, rs_d, rd_d, samples, n_dev, m, out.view(-1), lst)
    torch.cuda.synchronize()
    o, f = out.cpu().numpy(), full.cpu().numpy()
    assert np.array_equal(o[sel].view(np.int32), f[sel].view(np.int32))
    rest = np.setdiff1d(np.arange(m), sel)
    assert np.isnan(o[rest]).all()


def test_fused_field_over_a_shuffled_list(device):
    """A list in random order: nearly every sample of a 64-sample tile is a
    run of its own, so the fused field stages its corner rows in many rounds
    per tile; it still equals the split path (gather kernel + MLP kernel)
    bit for bit, slot by slot."""
    sc = shell()
    rs, rd = shell_rays(width=64)
    dsc = DeviceScene(sc, device)
    rs_d, rd_d = torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device)
    outs = []
    for split in (True, False):
        r = Renderer(device, precision="bf16", split_field=split)
        r.render(dsc, rs_d, rd_d, keep=True)
        torch.cuda.synchronize()
        m = int(r._ws["scalars"][1])
        samples = r._ws["c_samples"][: m * 4].view(m, 4).clone()
        perm = np.random.default_rng(11).permutation(m).astype(np.int32)
        idx = samples[:, 0].cpu().numpy()[perm]
        runs = 1 + (np.diff(idx.reshape(-1)[: (m // 64) * 64].reshape(-1, 64), axis=1) != 0).sum(1)
        assert runs.max() > 3 * 7  # past the third staging round
        n_dev = torch.tensor([m], dtype=torch.int32, device=device)
        out = torch.full((m, 4), float("nan"), dtype=torch.float32, device=device)
        r._field(dsc, rs_d, rd_d, samples, n_dev, m, out.view(-1), torch.from_numpy(perm).to(device))
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    assert not np.isnan(outs[1]).any()
    assert np.array_equal(outs[0].view(np.int32), outs[1].view(np.int32))
