"""Camera rays on the device (renderer.py:91-112, reader.py:92-119,
geometry.py:76-122): nsvf_generate_rays against the FP64 oracle, and a view
rendered from its camera (rays generated inside the traversal kernel)
against the same view rendered from the generated rays."""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell
from nsvf_amd import scene as S
from nsvf_amd._lib import NSVFError
from nsvf_amd.rays import ViewCamera, generate_rays, look_at_rotation
from nsvf_amd.render import DeviceScene, Renderer

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def cameras():
    c0 = ViewCamera.looking_at(S.sphere_camera(2.5, 11), 48, 64)              # 48 x 64
    pos = np.array([0.3, -2.1, 1.2])
    RT = np.eye(4)
    RT[:3, :3] = look_at_rotation(pos, up=(0.3, 0.2, -1.0))
    RT[:3, 3] = pos
    K = np.array([[90.0, 0, 61.3], [0, 85.0, 40.7], [0, 0, 1]])               # off-centre
    c1 = ViewCamera(K, RT, H=90, W=120, h=30, w=40)                            # downscaled grid
    return [c0, c1]


@pytest.mark.parametrize("k", [0, 1])
def test_generated_rays_match_fp64(device, k):
    cam = cameras()[k]
    rs, rd = generate_rays(cam, device)
    fx, fy, cx, cy = (cam.intrinsics[0, 0], cam.intrinsics[1, 1], cam.intrinsics[0, 2],
                      cam.intrinsics[1, 2])
    ref_s, ref_d = O.camera_rays(cam.H, cam.W, cam.h, cam.w, np.float32(fx), np.float32(fy),
                                 np.float32(cx), np.float32(cy),
                                 np.asarray(cam.inv_RT, np.float32).astype(np.float64))
    assert rs.shape == (cam.h * cam.w, 3)
    assert np.array_equal(N(rs), ref_s.astype(np.float32))
    err = np.abs(N(rd) - ref_d).max()
    print(f"camera {k}: max |d dir| vs FP64 {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_render_from_camera_equals_render_from_rays(device, precision):
    sc = shell()
    dsc = DeviceScene(sc, device)
    for cam in cameras():
        rs, rd = generate_rays(cam, device)
        r0 = Renderer(device, precision=precision, hierarchical=True)
        a = [N(t) for t in r0.render(dsc, rs, rd, width=cam.w)]
        r1 = Renderer(device, precision=precision, hierarchical=True)
        b = [N(t) for t in r1.render_camera(dsc, cam)]
        assert (a[2] < 1).any()                       # some pixels hit the shell
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_bad_cameras_raise(device):
    cam = cameras()[0]
    bad = ViewCamera(cam.intrinsics, cam.inv_RT, H=48, W=64, h=48, w=64)
    bad.w = 0
    with pytest.raises(NSVFError):
        generate_rays(bad, device)
