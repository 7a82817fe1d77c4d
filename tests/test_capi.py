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
