"""Voxel geometry helpers (fairnr/data/geometry.py), in torch.

Only the helpers the sparse-voxel path uses: lattice offsets, discretisation,
the octree wrapper, trilinear interpolation and image fill-in.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _ext


def offset_points(point_xyz, quarter_voxel=1.0, offset_only=False, bits=2):
    """geometry.py:229-238.  bits^3 lattice offsets (2i+1-bits)/(bits-1) *
    quarter_voxel around each point, x slowest and z fastest (meshgrid 'ij');
    with bits=2 these are the 8 corners in the order 4x + 2y + z."""
    c = torch.arange(1, 2 * bits, 2, device=point_xyz.device)
    ox, oy, oz = torch.meshgrid([c, c, c], indexing="ij")
    offset = (torch.cat([ox.reshape(-1, 1), oy.reshape(-1, 1), oz.reshape(-1, 1)], 1)
              .type_as(point_xyz) - bits) / float(bits - 1)
    if not offset_only:
        return point_xyz.unsqueeze(1) + offset.unsqueeze(0).type_as(point_xyz) * quarter_voxel
    return offset.type_as(point_xyz) * quarter_voxel


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


def trilinear_interp(p, q, point_feats):
    """geometry.py:195-200: weights prod(p*q + (1-p)(1-q)) over the 8 corners."""
    weights = (p * q + (1 - p) * (1 - q)).prod(dim=-1, keepdim=True)
    if point_feats.dim() == 2:
        point_feats = point_feats.view(point_feats.size(0), 8, -1)
    return (weights * point_feats).sum(1)


def fill_in(shape, hits, input, initial=1.0):
    """geometry.py:303-317: scatter the hit rays' values back to the frame."""
    input_sizes = list(input.size())
    if len(input_sizes) == len(shape) and all(shape[i] == input_sizes[i] for i in range(len(shape))):
        return input
    if isinstance(initial, torch.Tensor):
        output = initial.expand(*shape)
    else:
        output = input.new_ones(*shape) * initial
    if len(shape) == 1:
        return output.masked_scatter(hits, input)
    return output.masked_scatter(hits.unsqueeze(-1).expand(*shape), input)


def voxel_corner_keys(points, voxel_size):
    """encoder.py:269-276: unique sorted corner keys and each voxel's 8 corner
    ids (half-voxel integer units)."""
    coords, _ = discretize_points(points, voxel_size * 0.5)
    keys0 = offset_points(coords, 1.0).reshape(-1, 3)
    keys, feats = torch.unique(keys0, dim=0, sorted=True, return_inverse=True)
    return keys.long(), feats.reshape(-1, 8).long()
