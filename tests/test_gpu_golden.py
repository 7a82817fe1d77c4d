"""The fused pipeline on the GPU against the committed fixtures of the shell
scene's 64x64 view (tests/golden/shell64.npz): hit flags, the fixtures'
sorted hit rows and coarse sample streams bit for bit; the FP32 field's
image within 1e-4 of the fixtures' FP64 image end to end (its fine samples
follow its own coarse probabilities), the split-BF16 field's within 1e-3."""
import os

import numpy as np
import pytest
import torch

from nsvf_amd import scene as S
from nsvf_amd.render import DeviceScene, Renderer

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-4), ("bf16x3", 1e-3)])
def test_gpu_matches_the_fixtures(device, precision, tol):
    want = np.load(os.path.join(HERE, "golden", "shell64.npz"))
    sc = S.shell_scene()
    rs, rd = S.camera_rays(64, 64, S.sphere_camera(2.5, 7))
    dsc = DeviceScene(sc, device)
    r = Renderer(device, precision=precision, hierarchical=True, compat=True)
    rgb, depth, missed = (t.cpu().numpy() for t in r.render(
        dsc, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device), keep=True))
    torch.cuda.synchronize()
    k = r.kept
    assert np.array_equal(k["flag"].cpu().numpy().astype(bool), want["hit"])
    rays = want["rays"]
    W = want["idx"].shape[1]
    rows = k["hits"].cpu().numpy()[rays, :W]
    cnt = k["count"].cpu().numpy()[rays]
    m = np.arange(W)[None] < cnt[:, None]
    assert np.array_equal(np.where(m, rows[:, :, 2], -1), want["idx"])
    assert np.array_equal(np.where(m, rows[:, :, 0].view(np.float32), np.float32(1e4)), want["lo"])
    assert np.array_equal(np.where(m, rows[:, :, 1].view(np.float32), np.float32(1e4)), want["hi"])
    hl = k["hit_list"].cpu().numpy()
    pos = np.searchsorted(hl, rays)
    c_cnt, c_off = k["c_cnt"].cpu().numpy()[pos], k["c_off"].cpu().numpy()[pos]
    cols = [c.cpu().numpy() for c in k["c_samples"][:3]]
    Wc = want["c_idx"].shape[1]
    assert np.array_equal(c_cnt, (want["c_idx"] != -1).sum(1))
    for n in range(len(rays)):
        a, c = c_off[n], c_cnt[n]
        assert np.array_equal(cols[0][a:a + c], want["c_idx"][n, :c])
        assert np.array_equal(cols[1][a:a + c], want["c_depth"][n, :c])
        assert np.array_equal(cols[2][a:a + c], want["c_dist"][n, :c])
    assert Wc >= c_cnt.max()
    e = max(np.abs(rgb - want["rgb"]).max(), np.abs(missed - want["missed"]).max())
    print(f"{precision}: max |d rgb / alpha| vs the fixture {e:.2e}, "
          f"depth {np.abs(depth - want['depth']).max():.2e}")
    assert e < tol
