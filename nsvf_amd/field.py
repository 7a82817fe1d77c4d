"""Field MLP weights in the fused kernels' operand layout.

The field kernels keep activations in registers between layers by running
each layer as Y^T = W X^T on 16x16 MFMA tiles; that works when each layer's
weights are stored with their input features permuted to match the lanes'
accumulator rows (see csrc/field.hip).  This module does that packing once per
scene:

  BF16 (v_mfma_f32_16x16x32_bf16): fragment [m-tile][k-tile][lane][8]; lane
      l = 16g + m holds W[16 mt + m][32 kt + perm(g, i)], perm(g, i) = 4g + i
      for i < 4 and 16 + 4g + (i - 4) for i >= 4.
  FP32 (v_mfma_f32_16x16x4_f32): fragment [m-tile][k-tile][lane]; k-tile
      kt = 4t + i, lane l = 16g + m holds W[16 mt + m][16 t + 4g + i].

Rows are padded to 16 (the 4-wide head becomes one m-tile) and input features
to the k-tile size with zeros.  Biases: hidden layers back to back, then the
head padded to 16.
"""
from __future__ import annotations

import numpy as np
import torch

from .scene import FieldMLP


def _pad(W, rows, cols):
    out = np.zeros((rows, cols), np.float32)
    out[: W.shape[0], : W.shape[1]] = W
    return out


def bf16_bits(x):
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_round(x):
    """float32 values rounded to bfloat16 and back."""
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


def _perm_bf16():
    g = np.arange(4)[:, None]
    i = np.arange(8)[None, :]
    return np.where(i < 4, 4 * g + i, 16 + 4 * g + (i - 4))  # [4,8]


def pack_layer_bf16(W, kpad):
    O, K = W.shape
    Wp = _pad(W, ((O + 15) // 16) * 16, kpad)
    MT, KT = Wp.shape[0] // 16, kpad // 32
    lane = np.arange(64)
    m, g = lane % 16, lane // 16
    perm = _perm_bf16()[g]                                  # [64,8]
    rows = 16 * np.arange(MT)[:, None, None, None] + m[None, None, :, None]
    cols = 32 * np.arange(KT)[None, :, None, None] + perm[None, None, :, :]
    return Wp[rows, cols]                                   # [MT,KT,64,8]


def pack_layer_f32(W, kpad):
    O, K = W.shape
    Wp = _pad(W, ((O + 15) // 16) * 16, kpad)
    MT, KT = Wp.shape[0] // 16, kpad // 4
    lane = np.arange(64)
    m, g = lane % 16, lane // 16
    kt = np.arange(KT)
    rows = 16 * np.arange(MT)[:, None, None] + m[None, None, :]
    cols = 16 * (kt // 4)[None, :, None] + 4 * g[None, None, :] + (kt % 4)[None, :, None]
    return Wp[rows, cols]                                   # [MT,KT,64]


def split_bf16(x):
    """float32 x as bf16 bit patterns (hi, lo) with hi = bf16(x) and
    lo = bf16(x - hi) (the difference is exact in float32)."""
    x = np.ascontiguousarray(x, np.float32)
    hi = bf16_round(x)
    return bf16_bits(hi), bf16_bits((x - hi).astype(np.float32))


def pack_mlp(mlp: FieldMLP, device=None):
    """Returns dict with 'bf16' (uint16 fragments as int16 tensor), 'bf16x3'
    (the split-BF16 kernel's fragment pairs: per fragment its hi parts, then
    its lo parts), 'f32' (float fragments), 'bias' (float), and the shape
    (D, H, L)."""
    D = mlp.weights[0].shape[1]
    H = mlp.weights[0].shape[0]
    L = len(mlp.weights) - 1
    for w in mlp.weights[1:-1]:
        assert w.shape == (H, H), "hidden layers must be H x H"
    assert mlp.weights[-1].shape == (4, H)
    frags16, frags32 = [], []
    for li, W in enumerate(mlp.weights):
        K = W.shape[1]
        frags16.append(pack_layer_bf16(W, ((K + 31) // 32) * 32).reshape(-1))
        frags32.append(pack_layer_f32(W, ((K + 15) // 16) * 16).reshape(-1))
    bias = np.concatenate([b for b in mlp.biases[:-1]] + [_pad(mlp.biases[-1][None], 1, 16)[0]])
    all16 = np.concatenate(frags16)
    w16 = bf16_bits(all16).view(np.int16)
    hi, lo = split_bf16(all16)
    pairs = np.stack([hi.reshape(-1, 64 * 8), lo.reshape(-1, 64 * 8)], 1)   # [frag, 2, 512]
    w32 = np.concatenate(frags32).astype(np.float32)
    out = {
        "bf16": torch.from_numpy(w16.copy()),
        "bf16x3": torch.from_numpy(pairs.reshape(-1).view(np.int16).copy()),
        "f32": torch.from_numpy(w32),
        "bias": torch.from_numpy(bias.astype(np.float32)),
        "shape": (D, H, L),
    }
    if device is not None:
        for k in ("bf16", "bf16x3", "f32", "bias"):
            out[k] = out[k].to(device)
    return out
