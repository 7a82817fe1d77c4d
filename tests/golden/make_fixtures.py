"""Writes tests/golden/shell64.npz: regression fixtures of the oracle on the
spherical-shell scene (6,568 voxels, 32-d embeddings, field 3 x 128), its
64x64 camera (seed 7, radius 2.5):

  hit      [4096] bool     rays that hit a voxel
  rays     [128]  int64    a seeded subset of the hit rays (with the longest rows)
  idx/lo/hi [128, W]       their sorted hit rows, W the longest (-1 / 1e4 pads)
  c_idx/c_depth/c_dist [128, W]  their coarse inverse-CDF samples (reference
                           sampler mode, deterministic noise 0.5)
  rgb [4096, 3], depth [4096], missed [4096]  the hierarchical view (64 fine
                           samples) with the field and compositing in float64

Run from the repository root: python tests/golden/make_fixtures.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import nsvf_oracle as O  # noqa: E402
from nsvf_amd import scene as S  # noqa: E402


def fixtures():
    sc = S.shell_scene()
    rs, rd = S.camera_rays(64, 64, S.sphere_camera(2.5, 7))
    ref = O.render_view(sc, rs, rd, compat=True, dtype=np.float64)
    hit = ref["hits"]
    hit_rays = np.nonzero(hit)[0]
    counts = (ref["sorted_idx"] != -1).sum(1)
    longest = hit_rays[np.argsort(-counts[hit_rays], kind="stable")[:16]]
    rng = np.random.default_rng(3)
    rays = np.union1d(longest, rng.choice(hit_rays, 128 - len(longest), replace=False))
    while len(rays) < 128:
        rays = np.union1d(rays, rng.choice(hit_rays, 128 - len(rays), replace=False))
    W = int(counts.max())
    pos = np.searchsorted(hit_rays, rays)
    c_idx = ref["c_idx"][pos]
    cw = int((c_idx != -1).sum(1).max())
    return {"hit": hit, "rays": rays.astype(np.int64),
            "idx": ref["sorted_idx"][rays, :W], "lo": ref["sorted_lo"][rays, :W],
            "hi": ref["sorted_hi"][rays, :W],
            "c_idx": c_idx[:, :cw], "c_depth": ref["c_depth"][pos, :cw],
            "c_dist": ref["c_dist"][pos, :cw],
            "rgb": ref["rgb"], "depth": ref["depth"], "missed": ref["missed"]}


if __name__ == "__main__":
    out = os.path.join(HERE, "shell64.npz")
    np.savez_compressed(out, **fixtures())
    print(out, os.path.getsize(out), "bytes")
