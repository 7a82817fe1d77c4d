"""Scene files (encoder.py:212-301, 412-454, 1053-1059): bounding-box grids,
voxel .ply files (vertex and mesh forms), old-style .txt points and the
encoder's checkpoint buffers, on the CPU."""
import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell
from nsvf_amd import scene_io as IO


def test_bbox2voxels_order_and_spacing():
    """np.meshgrid's default indexing: z fastest, then x, then y; the box is
    inclusive (rounded step count + 1)."""
    pts = IO.bbox2voxels([0, 0, 0, 2, 1, 3, 1.0], 1.0)
    assert pts.dtype == np.float32 and pts.shape == (3 * 2 * 4, 3)
    assert np.array_equal(pts[0], [0, 0, 0]) and np.array_equal(pts[1], [0, 0, 1])
    assert np.array_equal(pts[4], [1, 0, 0]) and np.array_equal(pts[12], [0, 1, 0])
    assert np.array_equal(pts[-1], [2, 1, 3])
    assert np.unique(pts, axis=0).shape[0] == pts.shape[0]


def test_bbox_file(tmp_path):
    p = tmp_path / "bbox.txt"
    p.write_text("-0.5 -0.5 -0.25 0.5 0.5 0.25 0.25\n")
    pts, vs, index = IO.load_voxels(str(p))
    assert vs == 0.25 and index is None
    assert pts.shape == (5 * 5 * 3, 3) and pts.min() == np.float32(-0.5) and pts.max() == np.float32(0.5)


@pytest.mark.parametrize("text", [False, True])
def test_ply_round_trip(tmp_path, text):
    rng = np.random.default_rng(0)
    xyz = rng.standard_normal((17, 3)).astype(np.float32)
    q = np.arange(17, dtype=np.float32)
    faces = rng.integers(0, 17, (5, 4)).astype(np.int32)
    p = tmp_path / "v.ply"
    IO.write_ply(str(p), {"vertex": [("x", xyz[:, 0]), ("y", xyz[:, 1]), ("z", xyz[:, 2]),
                                     ("quality", q)],
                          "face": [("vertex_indices", faces)]}, text=text)
    head = p.read_bytes().split(b"end_header\n")[0].decode()
    assert head == ("ply\nformat " + ("ascii" if text else "binary_little_endian") + " 1.0\n"
                    "element vertex 17\nproperty float x\nproperty float y\nproperty float z\n"
                    "property float quality\nelement face 5\n"
                    "property list uchar int vertex_indices\n")
    d = IO.read_ply(str(p))
    assert np.array_equal(np.stack([d["vertex"]["x"], d["vertex"]["y"], d["vertex"]["z"]], 1), xyz)
    assert np.array_equal(d["vertex"]["quality"], q)
    assert np.array_equal(d["face"]["vertex_indices"], faces)


def test_export_vertices_keeps_original_indices(tmp_path):
    sc = shell()
    keep = np.arange(sc.num_voxels) % 3 != 0
    p = tmp_path / "voxels.ply"
    IO.export_voxels(sc, str(p), keep=keep)
    pts, vs, index = IO.load_voxels(str(p), voxel_size=sc.voxel_size)
    assert np.array_equal(pts, sc.points[keep]) and np.array_equal(index, np.nonzero(keep)[0])


def test_export_mesh_two_voxels(tmp_path):
    """Two voxels sharing a face: 12 distinct corners, 11 distinct faces; the
    loader takes the voxel size from the first face and every corner as a
    point (encoder.py:229-237)."""
    vs = 0.5
    sc = IO.scene_from_points(np.array([[0.25, 0.25, 0.25], [0.75, 0.25, 0.25]], np.float32), vs)
    p = tmp_path / "mesh.ply"
    IO.export_voxels(sc, str(p), return_mesh=True)
    d = IO.read_ply(str(p))
    assert d["vertex"]["x"].shape == (12,) and d["face"]["vertex_indices"].shape == (11, 4)
    pts, got_vs, index = IO.load_voxels(str(p))
    assert got_vs == vs and pts.shape == (12, 3)
    assert index is None   # the mesh's points are its corners, not one row per voxel
    assert set(map(tuple, pts.tolist())) == {(x, y, z) for x in (0.0, 0.5, 1.0)
                                             for y in (0.0, 0.5) for z in (0.0, 0.5)}


def test_txt_points(tmp_path):
    p = tmp_path / "points.txt"
    rows = np.array([[0, 0, 0, 0.1, 0.2, 0.3], [0, 0, 0, 0.5, 0.2, 0.3]])
    np.savetxt(p, rows)
    pts, vs, _ = IO.load_voxels(str(p), voxel_size=0.4)
    assert vs == 0.4 and np.array_equal(pts, rows[:, 3:].astype(np.float32))
    with pytest.raises(ValueError):
        IO.load_voxels(str(p))


def test_keys_match_the_oracle():
    sc = shell()
    s2 = IO.scene_from_points(sc.points, sc.voxel_size, embed_dim=8)
    keys, feats = O.voxel_keys(sc.points, sc.voxel_size)
    assert np.array_equal(s2.keys, keys) and np.array_equal(s2.feats, feats)
    assert s2.values.shape == (keys.shape[0], 8)


def test_checkpoint_buffers_and_round_trip(tmp_path):
    sc = shell()
    keep = (np.arange(sc.num_voxels) % 5 != 1).astype(np.int64)
    p = tmp_path / "ckpt.pt"
    IO.save_checkpoint(sc, str(p), keep=keep)
    sd = torch.load(str(p), weights_only=True)["model"]
    want = {"points": (torch.float32, (sc.num_voxels, 3)), "keys": (torch.int64, sc.keys.shape),
            "feats": (torch.int64, (sc.num_voxels, 8)), "num_keys": (torch.int64, ()),
            "keep": (torch.int64, (sc.num_voxels,)), "voxel_size": (torch.float32, ()),
            "step_size": (torch.float32, ()), "max_hits": (torch.float32, ()),
            "values.weight": (torch.float32, sc.values.shape)}
    for k, (dt, shape) in want.items():
        t = sd["encoder." + k]
        assert t.dtype == dt and tuple(t.shape) == tuple(shape), k
    s2 = IO.load_checkpoint(str(p), mlp=sc.mlp)
    kb = keep.astype(bool)
    assert np.array_equal(s2.points, sc.points[kb]) and np.array_equal(s2.feats, sc.feats[kb])
    assert np.array_equal(s2.values, sc.values) and np.array_equal(s2.keys, sc.keys)
    assert (s2.voxel_size, s2.step_size, s2.max_hits) == (sc.voxel_size, sc.step_size, sc.max_hits)
    perm = np.random.default_rng(1).permutation(sc.num_voxels)
    s3 = IO.load_checkpoint(sd, mlp=sc.mlp, voxel_index=perm)
    assert np.array_equal(s3.points, sc.points[perm][kb[perm]])
