"""Camera rays on the device: the reference's ray generation
(fairnr/renderer.py:91-112 ``generate_rays``, fairnr/modules/reader.py:92-119
``ImageReader.forward``, fairnr/data/geometry.py:68-122, data_utils.py:57-65)
with the same names and argument meaning, computed by ``nsvf_generate_rays``
(or inside the traversal by ``Renderer.render_camera``)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import Camera as _CCamera, check, lib


def get_uv(H, W, h, w):
    """data_utils.py:57-65: pixel coordinates [2, h, w] (u = column x W/w,
    v = row x H/h, float32) and the two scale factors."""
    uv = np.flip(np.mgrid[0:h, 0:w], axis=0).astype(np.float32)
    uv[0] = uv[0] * np.float32(W / w)
    uv[1] = uv[1] * np.float32(H / h)
    return uv, [float(H / h), float(W / w)]


def parse_intrinsics(intrinsics):
    """geometry.py:68-73: fx, fy, cx, cy of a 3x3 / 4x4 intrinsics matrix."""
    k = np.asarray(intrinsics, np.float64)
    return float(k[0, 0]), float(k[1, 1]), float(k[0, 2]), float(k[1, 2])


def look_at_rotation(camera_position, at=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0)):
    """geometry.py:125-160 (OpenCV convention, the renderer's call): the
    camera-to-world rotation with columns x, y, z, z towards `at`."""
    c = np.asarray(camera_position, np.float64)

    def unit(v):
        return v / (np.linalg.norm(v) + 1e-8)
    z = unit(np.asarray(at, np.float64) - c)
    x = unit(np.cross(np.asarray(up, np.float64), z))
    y = unit(np.cross(z, x))
    return np.stack([x, y, z], 1)


@dataclass
class ViewCamera:
    """One view: intrinsics of the full H x W image, the camera-to-world
    pose inv_RT (4x4 or 3x4), and the rendered grid h x w."""
    intrinsics: np.ndarray
    inv_RT: np.ndarray
    H: int
    W: int
    h: int = 0
    w: int = 0

    def __post_init__(self):
        self.h = int(self.h or self.H)
        self.w = int(self.w or self.W)
        self.H, self.W = int(self.H), int(self.W)

    @property
    def num_rays(self):
        return self.h * self.w

    def c(self):
        fx, fy, cx, cy = parse_intrinsics(self.intrinsics)
        rt = np.asarray(self.inv_RT, np.float32)[:3, :4]
        cam = _CCamera(fx, fy, cx, cy, (ctypes.c_float * 12)(*rt.reshape(-1).tolist()),
                       self.H, self.W, self.h, self.w)
        return cam

    @staticmethod
    def looking_at(cam_pos, H, W, focal=None, at=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0)):
        """A camera at `cam_pos` looking at `at` (the renderer's path
        generator, renderer.py:92-100), focal 1.2 W and the principal point
        at the image centre unless given."""
        f = 1.2 * W if focal is None else focal
        K = np.array([[f, 0, W / 2.0], [0, f, H / 2.0], [0, 0, 1]], np.float64)
        RT = np.eye(4)
        RT[:3, :3] = look_at_rotation(cam_pos, at, up)
        RT[:3, 3] = cam_pos
        return ViewCamera(K, RT, H, W)


def generate_rays(camera: ViewCamera, device):
    """Rays of `camera` on `device`: (ray_start [h*w,3], ray_dir [h*w,3])
    float32, row-major pixel order."""
    R = camera.num_rays
    rs = torch.empty(R, 3, dtype=torch.float32, device=device)
    rd = torch.empty(R, 3, dtype=torch.float32, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(torch.device(device)).cuda_stream)
    check(lib().nsvf_generate_rays(ctypes.byref(camera.c()), ctypes.c_void_p(rs.data_ptr()),
                                   ctypes.c_void_p(rd.data_ptr()), st), "generate_rays")
    return rs, rd
