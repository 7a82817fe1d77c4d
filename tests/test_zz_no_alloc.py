"""The allocation-free field entry point (nsvf_field_eval_static) against
nsvf_field_eval with a caller workspace, through the C ABI: the same bits on
two calls in a row -- the second with a smaller input (and an evaluation
list) -- for split BF16 and BF16 through the static scratch, and FP32; sizes
beyond the scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from _cases import shell, shell_rays
from nsvf_amd import _lib
from nsvf_amd.render import DeviceScene, Renderer, PRECISIONS, _WEIGHTS

pytestmark = pytest.mark.gpu


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@pytest.fixture(scope="module")
def stream_case(device):
    sc = shell()
    rs, rd = shell_rays(width=64)
    dsc = DeviceScene(sc, device)
    r = Renderer(device, precision="fp32", hierarchical=False)
    rs_d, rd_d = torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device)
    r.render(dsc, rs_d, rd_d, keep=True)
    torch.cuda.synchronize()
    m = int(r._ws["scalars"][1])
    samples = r._ws["c_samples"][: m * 4].view(m, 4).clone()
    return dsc, rs_d, rd_d, samples, m


def _eval(L, dsc, rs, rd, samples, m, prec, lst, static):
    """out [m, 4] from one call (NaN where the call writes nothing)."""
    dev = samples.device
    n = len(lst) if lst is not None else m
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    out = torch.full((m, 4), float("nan"), dtype=torch.float32, device=dev)
    w = dsc.mlp[_WEIGHTS[PRECISIONS[prec]]]
    args_head = (PRECISIONS[prec], dsc.D, dsc.H, dsc.L, P(dsc.centers), P(dsc.feats),
                 P(dsc.values), dsc.voxel_size, P(rs), P(rd), P(samples), P(lst), P(n_dev), m,
                 P(w), P(dsc.mlp["bias"]), P(out))
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if static:
        rc = L.nsvf_field_eval_static(*args_head, st)
    else:
        nb = int(L.nsvf_field_workspace_bytes(PRECISIONS[prec], dsc.D, dsc.H, dsc.L, m))
        ws = torch.empty(max(nb, 16), dtype=torch.uint8, device=dev)
        rc = L.nsvf_field_eval(*args_head, P(ws) if nb > 0 else None, nb, st)
    assert rc == 0, L.nsvf_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("prec", ["bf16x3", "bf16", "fp32"])
def test_static_scratch_equals_workspace_twice(stream_case, prec):
    L = _lib.lib()
    dsc, rs, rd, samples, m = stream_case
    print(f"{prec}: first call, {m} samples", flush=True)
    a = _eval(L, dsc, rs, rd, samples, m, prec, None, True)
    b = _eval(L, dsc, rs, rd, samples, m, prec, None, False)
    assert not np.isnan(a).any()
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    # second call: a smaller stream, every third slot through a list
    m2 = m // 3
    lst = torch.arange(0, m2, 3, dtype=torch.int32, device=samples.device)
    small = samples[:m2].clone()
    print(f"{prec}: second call, {m2} samples, list of {len(lst)}", flush=True)
    a2 = _eval(L, dsc, rs, rd, small, m2, prec, lst, True)
    b2 = _eval(L, dsc, rs, rd, small, m2, prec, lst, False)
    sel = np.arange(0, m2, 3)
    assert np.array_equal(a2[sel].view(np.int32), b2[sel].view(np.int32))
    assert np.array_equal(a2[sel].view(np.int32), a[sel].view(np.int32))
    rest = np.setdiff1d(np.arange(m2), sel)
    assert np.isnan(a2[rest]).all()


def test_static_scratch_refuses_larger_streams(stream_case):
    L = _lib.lib()
    dsc, rs, rd, samples, m = stream_case
    cap = int(L.nsvf_field_static_scratch_bytes())
    assert cap == 64 << 20
    too_many = cap // 128 + 64       # split BF16 at D = 32: 128 bytes per sample
    n_dev = torch.tensor([0], dtype=torch.int32, device=samples.device)
    rc = L.nsvf_field_eval_static(2, dsc.D, dsc.H, dsc.L, P(dsc.centers), P(dsc.feats),
                                  P(dsc.values), dsc.voxel_size, P(rs), P(rd), P(samples), None,
                                  P(n_dev), too_many, P(dsc.mlp["bf16x3"]), P(dsc.mlp["bias"]),
                                  P(samples), None)
    assert rc == -1 and b"static scratch" in L.nsvf_last_error()
