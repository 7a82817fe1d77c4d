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


def test_scene_files_render_as_the_oracle(device, tmp_path):
    """Scenes read from the reference's voxel files render as the oracle
    renders the same scene: a bbox.txt grid (encoder.py:215-231, 1053-1059)
    and its vertex .ply export read back (encoder.py:250-267, 412-454) --
    hit flags, hit rows and the coarse sample stream bit-exact, the FP32
    image within 1e-5 of the FP64 oracle on the same samples."""
    import nsvf_oracle as O
    from nsvf_amd import scene as S
    box = tmp_path / "bbox.txt"
    box.write_text("-0.4 -0.3 -0.35 0.4 0.3 0.35 0.1\n")
    pts, vs, _ = IO.load_voxels(str(box))
    assert len(pts) == 9 * 7 * 8
    sc = IO.scene_from_points(pts, vs, seed=3)
    ply = tmp_path / "voxels.ply"
    IO.export_voxels(sc, str(ply))
    pts2, vs2, index = IO.load_voxels(str(ply), voxel_size=vs)
    assert np.array_equal(pts2, pts) and np.array_equal(index, np.arange(len(pts)))
    sc2 = IO.scene_from_points(pts2, vs2, values=sc.values, mlp=sc.mlp)
    rs, rd = S.camera_rays(40, 40, S.sphere_camera(1.6, 4))
    ref = O.render_view(sc, rs, rd, hierarchical=False, compat=True)
    assert ref["hits"].any() and not ref["hits"].all()
    for scene in (sc, sc2):
        r = Renderer(device, precision="fp32", hierarchical=False)
        rgb, depth, missed = r.render(DeviceScene(scene, device), torch.from_numpy(rs).to(device),
                                      torch.from_numpy(rd).to(device), keep=True)
        torch.cuda.synchronize()
        k = r.kept
        hits = N(k["flag"]).astype(bool)
        assert np.array_equal(hits, ref["hits"])
        cnt = N(k["count"])
        rows = N(k["hits"])
        for ray in np.nonzero(hits)[0]:
            n = cnt[ray]
            assert np.array_equal(rows[ray, :n, 2], ref["sorted_idx"][ray, :n])
        c_cnt, c_off = N(k["c_cnt"]), N(k["c_off"])
        ci = ref["c_idx"]
        assert np.array_equal(c_cnt, (ci != -1).sum(1))
        recs = N(k["c_samples"][0])
        for n in range(len(c_cnt)):
            assert np.array_equal(recs[c_off[n]: c_off[n] + c_cnt[n]], ci[n, : c_cnt[n]])
        assert np.abs(N(rgb) - ref["rgb"]).max() < 1e-5
        assert np.abs(N(missed) - ref["missed"]).max() < 1e-5
        assert np.abs(N(depth) - ref["depth"]).max() < 1e-5
