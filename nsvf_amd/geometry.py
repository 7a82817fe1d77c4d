"""Voxel geometry helpers (fairnr/data/geometry.py), in torch: the
discretisation of voxel centres and the octree wrapper the device scene is
built with.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _ext


def discretize_points(voxel_points, voxel_size):
    """geometry.py:241-247.  Integer coordinates relative to the minimum point
    and the mean residual.  The mean is taken in float64 and rounded to
    float32 (the reference leaves the float32 summation order to torch)."""
    minimal = voxel_points.min(dim=0, keepdim=True)[0]
    idx = ((voxel_points - minimal) / voxel_size).round_().long()
    residual = (voxel_points.double() - (idx.type_as(voxel_points) * voxel_size).double()).mean(
        0, keepdim=True).float()
    return idx, residual


def octree_depth(coords):
    """geometry.py:322-323: ceil(log2(max coordinate range)) - 1, at least 0
    (a single voxel, where the reference's log2(0) is undefined)."""
    ranges = coords.max(0)[0] - coords.min(0)[0]
    r = int(ranges.max().item())
    if r <= 0:
        return 0
    return max(int(math.ceil(math.log2(r))) - 1, 0)


def build_easy_octree(points, half_voxel):
    """geometry.py:320-327.  Returns node centres [T,3] f32 (on points'
    device) and children [T,9] i32.  The centre is (max + min) / 2 of the
    integer coordinates, which is integral for voxel centres (their
    coordinates are all even)."""
    coords, residual = discretize_points(points, half_voxel)
    depth = octree_depth(coords)
    center = (coords.max(0)[0] + coords.min(0)[0]) // 2
    centers, children = _ext.build_octree(center.cpu(), coords.cpu(), depth)
    centers = centers.to(points.device).float() * half_voxel + residual
    return centers, children.to(points.device)
