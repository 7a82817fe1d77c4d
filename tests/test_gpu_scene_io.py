"""A scene saved as the encoder's checkpoint buffers and loaded back renders
bit-identically (and a checkpoint with pruned voxels renders as the pruned
scene)."""
import numpy as np
import pytest
import torch

from _cases import shell, shell_rays
from nsvf_amd import scene_io as IO
from nsvf_amd.render import DeviceScene, Renderer
from nsvf_amd.voxels import prune_scene

pytestmark = pytest.mark.gpu


def N(t):
    return t.detach().cpu().numpy()


def view(sc, device):
    rs, rd = shell_rays(64, seed=7)
    r = Renderer(device, precision="fp32", hierarchical=True)
    return [N(t) for t in r.render(DeviceScene(sc, device), torch.from_numpy(rs).to(device),
                                   torch.from_numpy(rd).to(device), width=64)]


def test_checkpoint_round_trip_renders_the_same(device, tmp_path):
    sc = shell()
    keep = (np.arange(sc.num_voxels) % 4 != 0).astype(np.int64)
    for k in (np.ones(sc.num_voxels, np.int64), keep):
        p = tmp_path / "ckpt.pt"
        IO.save_checkpoint(sc, str(p), keep=k)
        got = view(IO.load_checkpoint(str(p), mlp=sc.mlp), device)
        ref = view(prune_scene(sc, k.astype(bool)), device)
        assert (ref[2] < 1).any()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
