"""The C-ABI library loads, exports every declared entry point, and its host
code (the octree builder) matches the oracle.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import nsvf_oracle as O
from _cases import shell
from nsvf_amd import _ext, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nsvf_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nsvf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert L.nsvf_version() >= 100


def test_last_error_reports_bad_arguments():
    L = _lib.lib()
    rc = L.nsvf_aabb_intersect(None, None, None, 1, -1, 1, 1, 0.1, 4, None, None, None, None)
    assert rc == -1
    assert b"negative" in L.nsvf_last_error()


def _octree_both(coords):
    center = ((coords.max(0) + coords.min(0)) // 2).astype(np.int64)
    r = int((coords.max(0) - coords.min(0)).max())
    depth = max(int(np.ceil(np.log2(r))) - 1, 0) if r > 0 else 0
    ref = O.build_octree(center, coords, depth)
    got = _ext.build_octree(torch.from_numpy(center), torch.from_numpy(coords), depth)
    return ref, got


def test_build_octree_shell_matches_oracle():
    sc = shell()
    coords, _ = O.discretize_points(sc.points, np.float32(sc.voxel_size) / 2)
    (rc, rch), (gc, gch) = _octree_both(coords)
    assert rc.shape == (14521, 3)     # node count of this scene
    assert np.array_equal(gc.numpy(), rc) and np.array_equal(gch.numpy(), rch)


@pytest.mark.parametrize("seed", range(8))
def test_build_octree_random_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    g = int(rng.integers(2, 48))
    m = rng.random((g, g, g)) < rng.uniform(0.01, 0.6)
    ijk = np.argwhere(m)
    if len(ijk) < 2:
        ijk = np.array([[0, 0, 0], [1, 1, 1]])
    ijk = ijk[rng.permutation(len(ijk))]
    coords = (ijk * 2).astype(np.int64)
    (rc, rch), (gc, gch) = _octree_both(coords)
    assert np.array_equal(gc.numpy(), rc) and np.array_equal(gch.numpy(), rch)


def test_build_octree_rejects_shared_leaf_slot():
    coords = np.array([[0, 0, 0], [0, 0, 0], [4, 4, 4]], np.int64)
    with pytest.raises(RuntimeError):
        _ext.build_octree(torch.tensor([2, 2, 2]), torch.from_numpy(coords), 1)


def test_cpu_tensors_are_rejected_like_the_reference():
    z = torch.zeros(1, 2, 3)
    with pytest.raises(RuntimeError, match="CUDA"):
        _ext.aabb_intersect(z, z, z, 0.1, 4)
    with pytest.raises(RuntimeError, match="CUDA"):
        _ext.svo_intersect(z, z, z, torch.zeros(1, 2, 9, dtype=torch.int32), 0.1, 4)
    with pytest.raises(NotImplementedError):
        _ext.ball_intersect(z, z, z, 0.1, 4)


def test_render_hits_rejects_trees_deeper_than_the_traversal():
    """The fused traversal holds one record per level (depth + 2 levels, at
    most 24): a deeper tree is refused with NSVF_ESTACK before any launch
    instead of losing the hits below the last level."""
    L = _lib.lib()
    buf = (ctypes.c_float * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    args = lambda depth: (p, p, 64, 0, p, 3, depth, 60, 0.01, p, p, p, p, p, None)
    assert L.nsvf_render_hits(*args(23)) == -3
    assert b"depth 23" in L.nsvf_last_error()
    assert L.nsvf_render_hits(*args(-2)) == -1
    cam = _lib.Camera()
    cam.H = cam.W = cam.h = cam.w = 8
    rc = L.nsvf_render_hits_camera(ctypes.byref(cam), p, 3, 25, 60, 0.01, p, p, p, p, p, p, p,
                                   None)
    assert rc == -3


def test_tree_depth_of_packed_trees():
    """DeviceScene passes the depth the root's scale gives (octree.cpp:119):
    the shell scene's tree is 5 deep, a one-voxel tree 0, and a synthetic
    chain 25 deep is beyond the traversal."""
    from nsvf_amd.render import tree_depth
    sc = shell()
    coords, _ = O.discretize_points(sc.points, np.float32(sc.voxel_size) / 2)
    (_, rch), _ = _octree_both(coords)
    assert tree_depth(rch) == 5
    ch = np.full((2, 9), -1, np.int32)
    ch[0, 8], ch[1, :] = 1, [0, -1, -1, -1, -1, -1, -1, -1, 2]
    assert tree_depth(ch) == 0
    deep = np.full((27, 9), -1, np.int32)
    deep[0, 8] = 1
    for k in range(1, 27):
        deep[k, 0], deep[k, 8] = k - 1, 1 << k
    assert tree_depth(deep) == 25
    with pytest.raises(ValueError):
        tree_depth(np.array([[0, -1, -1, -1, -1, -1, -1, -1, 3]], np.int32))


def test_new_entry_points_reject_bad_arguments():
    """Argument checks that return before any device work (no GPU needed):
    the training MLP's built shape and pointers, the marching entry points' chunk
    and pointers, the samplers' 32-bit noise seeds and the split-BF16 tile
    writer's size, pointers, shape and workspace."""
    L = _lib.lib()
    assert L.nsvf_mlp_train_supported(32, 128, 3) == 1
    assert L.nsvf_mlp_train_supported(48, 256, 4) == 0
    assert L.nsvf_mlp_train_packed_floats(48, 256, 4) == -1
    assert L.nsvf_mlp_train_param_floats(32, 128, 3) == 128 * 32 + 128 + 2 * (128 * 128 + 128) + 4 * 128 + 4
    assert L.nsvf_mlp_train_partial_floats(32, 128, 3, 0) == L.nsvf_mlp_train_param_floats(32, 128, 3)
    assert L.nsvf_mlp_train_forward(48, 256, 4, None, None, 10, None, None) == -1
    assert b"shape" in L.nsvf_last_error()
    assert L.nsvf_mlp_train_forward(32, 128, 3, None, None, 10, None, None) == -1
    assert b"null" in L.nsvf_last_error()
    assert L.nsvf_mlp_train_forward(32, 128, 3, None, None, 0, None, None) == 0
    assert L.nsvf_mlp_train_backward(32, 128, 3, None, None, 10, None, None, None, 0, None, None) == -1
    # split-BF16 backward: 144 fragments (W_0 8, W_1 / W_2 32 each, and the
    # transposes 8, 32, 32); its fragments must be given and 16-byte aligned
    assert L.nsvf_mlp_train_x3_frags(32, 128, 3) == 144
    assert L.nsvf_mlp_train_x3_frags(48, 256, 4) == -1
    assert L.nsvf_mlp_train_backward_x3(32, 128, 3, None, None, None, 10, None, None, None, 0,
                                        None, None) == -1
    assert b"fragments" in L.nsvf_last_error()
    assert L.nsvf_mlp_train_backward_x3(32, 128, 3, None, 8, None, 10, None, None, None, 0,
                                        None, None) == -1
    assert b"aligned" in L.nsvf_last_error()
    assert L.nsvf_mlp_train_backward_x3(48, 256, 4, None, 16, None, 10, None, None, None, 0,
                                        None, None) == -1
    assert b"shape" in L.nsvf_last_error()
    assert L.nsvf_march_begin(None, 10, None, 0, 5, None, None, None, None) == -1
    assert b"chunk" in L.nsvf_last_error()
    assert L.nsvf_march_begin(None, 10, None, 4, 5, None, None, None, None) == -1
    assert b"null" in L.nsvf_last_error()
    assert L.nsvf_march_step(None, 10, None, None, None, None, 4, 5, 1.0, None, None, None, None,
                             None, None, None) == -1
    assert L.nsvf_march_end(None, 10, None, None, None, None, None) == -1
    assert L.nsvf_march_begin(None, 0, None, 0, 5, None, None, None, None) == 0
    assert L.nsvf_march_hist_ints(10, 4, 5) == 3 * 6
    rc = L.nsvf_render_sample_coarse(None, None, 10, None, None, None, None, None, 60, 0.1, 1, 64.0,
                                     None, None, None, None, None, 0, None, 1 << 32, None)
    assert rc == -1 and b"noise seed" in L.nsvf_last_error()
    # the split-BF16 operand tiles from given features
    need = L.nsvf_field_workspace_bytes(2, 32, 128, 3, 100)
    assert need == 2 * 128 * 4 * 16                     # hi and lo, samples padded to 64
    assert L.nsvf_field_tiles_bf16x3(32, 128, 3, None, 0, None, 0, None) == 0
    assert L.nsvf_field_tiles_bf16x3(32, 128, 3, None, -1, None, 0, None) == -1
    assert L.nsvf_field_tiles_bf16x3(32, 128, 3, None, 100, None, need, None) == -1
    assert b"null" in L.nsvf_last_error()
    assert L.nsvf_field_tiles_bf16x3(16, 64, 2, 1, 100, 1, need, None) == -1
    assert b"shape" in L.nsvf_last_error()
    assert L.nsvf_field_tiles_bf16x3(32, 128, 3, 1, 100, 1, need - 1, None) == -1
    assert b"workspace" in L.nsvf_last_error()
    # the ordered traversal: its order pointer is required
    assert L.nsvf_render_hits_ordered(1, 1, 10, None, 1, 5, 2, 60, 0.1, 1, 1, 1, 1, 1, None) == -1
    assert b"null" in L.nsvf_last_error()
    assert L.nsvf_render_hits_ordered(None, None, 0, None, 1, 5, 2, 60, 0.1, None, None, None,
                                      None, None, None) == 0
    assert L.nsvf_render_hits_ordered(1, 1, 10, 1, 1, 5, 25, 60, 0.1, 1, 1, 1, 1, 1, None) == -3
    assert L.nsvf_ray_order_keys(None, None, 10, None, None, None) == -1
    assert b"null" in L.nsvf_last_error()
    assert L.nsvf_ray_order_keys(None, None, 0, None, None, None) == 0
    assert L.nsvf_ray_order_keys(None, None, -1, None, None, None) == -1
