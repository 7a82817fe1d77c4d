"""Differentiable rendering for training: the gradient of an image loss with
respect to the corner embeddings and the field MLP.

The reference trains through torch autograd over its forward
(fairnr/criterions/rendering_loss.py:37-139 -> NSVFModel._forward): the
intersection and sampling ops are non-differentiable (their autograd
Functions mark every output so, clib/__init__.py:58-300), and gradients flow
through three parts:

  trilinear gather   encoder.py:582-590 -- F.embedding backward, an index-add
                     into the embedding table;
  field MLP          field.py:217-279 -- nn.Linear layers;
  compositing        renderer.py:119-121,195-218 -- relu, exp, cumsum.

Here the geometry stages are the fused renderer's kernels (one coarse sample
stream per batch of rays), and the gather, the MLP and the compositing are
autograd Functions over hand-written HIP forward and backward kernels, all
deterministic and atomic-free (the MLP on FP32 MFMA; torch.nn.Linear with
torch autograd, as in the reference, on request).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from ._lib import check, lib
from .field import pack_layer_bf16
from .render import BG_COLOR, BG_DEPTH, CoarseStream, DeviceScene, Renderer, coherent_order
from .scene import VoxelScene


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _need_cuda(t, name):
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")


class TrilinearGather(torch.autograd.Function):
    """emb [M,D] = sum_c w_c(x_m) values[feats[s_idx[m]][c]] (encoder.py:582-590);
    rows with s_idx -1 are zero.  Differentiable in `values` only, as in the
    reference (sample points come from non-differentiable sampling)."""

    @staticmethod
    def forward(ctx, values, s_idx, xyz, centers, feats, voxel_size, plan=None):
        for t, n in ((values, "values"), (s_idx, "s_idx"), (xyz, "xyz"), (centers, "centers"),
                     (feats, "feats")):
            _need_cuda(t, n)
        if values.dtype != torch.float32 or xyz.dtype != torch.float32:
            raise RuntimeError("values and xyz must be float32")
        if s_idx.dtype != torch.int32 or feats.dtype != torch.int32:
            raise RuntimeError("s_idx and feats must be int32")
        values = values.contiguous()
        s_idx, xyz = s_idx.contiguous(), xyz.contiguous()
        centers, feats = centers.contiguous(), feats.contiguous()
        M, (K, D) = s_idx.shape[0], values.shape
        out = torch.empty((M, D), dtype=torch.float32, device=values.device)
        check(lib().nsvf_trilinear_gather(_p(s_idx), _p(xyz), M, _p(centers), _p(feats),
                                          _p(values), D, float(voxel_size), _p(out),
                                          _stream(values.device)), "trilinear_gather")
        ctx.save_for_backward(s_idx, xyz, centers, feats)
        ctx.shape = (K, D)
        ctx.voxel_size = float(voxel_size)
        ctx.plan = plan
        return out

    @staticmethod
    def backward(ctx, grad_out):
        s_idx, xyz, centers, feats = ctx.saved_tensors
        K, D = ctx.shape
        grad_values = None
        if ctx.needs_input_grad[0]:
            grad_values = trilinear_gather_backward(s_idx, xyz, centers, feats, K, grad_out,
                                                    ctx.voxel_size, ctx.plan)
        return grad_values, None, None, None, None, None, None


def trilinear_gather(values, s_idx, xyz, centers, feats, voxel_size, plan=None):
    """plan: optional gather_backward_plan(feats, K) of the scene, reused
    by the backward instead of sorting the corner incidences again."""
    return TrilinearGather.apply(values, s_idx, xyz, centers, feats, voxel_size, plan)


def gather_backward_plan(feats, num_keys):
    """The corner-key order of a scene's (voxel, corner) incidences (a
    uint8 device buffer): what every backward of the gather over this scene
    would sort again."""
    L = lib()
    dev = feats.device
    N = feats.shape[0]
    st = _stream(dev)
    nb = int(L.nsvf_trilinear_gather_backward_plan_bytes(N, num_keys))
    ws_bytes = int(L.nsvf_trilinear_gather_backward_workspace(0, N, num_keys, 4, st))
    if nb < 0 or ws_bytes < 0:
        check(-1, "trilinear_gather_backward_plan_bytes")
    plan = torch.empty(nb, dtype=torch.uint8, device=dev)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    check(L.nsvf_trilinear_gather_backward_plan(_p(feats.contiguous()), N, num_keys, _p(plan), nb,
                                                _p(ws), ws_bytes, st),
          "trilinear_gather_backward_plan")
    return plan


def trilinear_gather_backward(s_idx, xyz, centers, feats, num_keys, grad_out, voxel_size,
                              plan=None):
    """grad_values [K,D] of the gather (deterministic; see the HIP source).
    plan: gather_backward_plan(feats, num_keys), or None to sort the corner
    incidences in this call (same result)."""
    L = lib()
    dev = grad_out.device
    grad_out = grad_out.contiguous().float()
    M, D = grad_out.shape
    N = centers.shape[0]
    st = _stream(dev)
    nbytes = int(L.nsvf_trilinear_gather_backward_workspace(M, N, num_keys, D, st))
    if nbytes < 0:
        check(-1, "trilinear_gather_backward_workspace")
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    grad_values = torch.empty((num_keys, D), dtype=torch.float32, device=dev)
    check(L.nsvf_trilinear_gather_backward_planned(_p(s_idx), _p(xyz), M, _p(centers), _p(feats),
                                                   N, num_keys, D, float(voxel_size),
                                                   _p(grad_out), _p(grad_values), _p(plan),
                                                   _p(ws), nbytes, st),
          "trilinear_gather_backward")
    return grad_values


class StreamComposite(torch.autograd.Function):
    """Front-to-back compositing of a sample stream (renderer.py:119-121,
    195-218).  field [slots,4] = (sigma, r, g, b) per slot; returns per-ray
    (rgb [R,3], depth [R], missed [R]) with rays that hit nothing at
    (0, 0, 1).  Differentiable in `field`."""

    @staticmethod
    def forward(ctx, field, g: CoarseStream):
        field = field.contiguous()
        if field.dtype != torch.float32 or field.shape != (g.slots, 4):
            raise RuntimeError(f"field must be float32 [{g.slots}, 4]")
        dev = field.device
        R = g.rays
        rgb = torch.zeros((R, 3), dtype=torch.float32, device=dev)
        depth = torch.zeros(R, dtype=torch.float32, device=dev)
        missed = torch.ones(R, dtype=torch.float32, device=dev)
        if g.n_hit > 0:
            check(lib().nsvf_render_composite(_p(g.hit_list), _p(g.scal[0:1]), g.n_hit, _p(g.off),
                                              _p(g.cnt), _p(g.samples), _p(field), None, _p(rgb),
                                              _p(depth), _p(missed), _stream(dev)), "composite")
        ctx.save_for_backward(field)
        ctx.stream = g
        return rgb, depth, missed

    @staticmethod
    def backward(ctx, g_rgb, g_depth, g_missed):
        (field,) = ctx.saved_tensors
        g = ctx.stream
        dev = field.device
        R = g.rays
        g_rgb = torch.zeros((R, 3), device=dev) if g_rgb is None else g_rgb.contiguous().float()
        g_depth = torch.zeros(R, device=dev) if g_depth is None else g_depth.contiguous().float()
        g_missed = torch.zeros(R, device=dev) if g_missed is None else g_missed.contiguous().float()
        grad_field = torch.zeros_like(field)
        if g.n_hit > 0:
            check(lib().nsvf_render_composite_backward(
                _p(g.hit_list), _p(g.scal[0:1]), g.n_hit, _p(g.off), _p(g.cnt), _p(g.samples),
                _p(field), _p(g_rgb), _p(g_depth), _p(g_missed), _p(grad_field), _stream(dev)),
                "composite_backward")
        return grad_field, None


def _mlp_shape(weights):
    return int(weights[0].shape[1]), int(weights[0].shape[0]), len(weights) - 1


def mlp_train_supported(D, H, L):
    """Whether the HIP training MLP is built for input width D, L hidden
    layers of width H (and the 4-wide head)."""
    return bool(lib().nsvf_mlp_train_supported(int(D), int(H), int(L)))


def pack_mlp_train(weights, biases):
    """The packed operand fragments and biases of the training MLP kernels
    (a flat float32 device tensor), from torch-layout weights [out, in]."""
    L_ = lib()
    D, H, L = _mlp_shape(weights)
    n = int(L_.nsvf_mlp_train_packed_floats(D, H, L))
    if n < 0:
        raise RuntimeError(f"training MLP: shape D={D} H={H} L={L} is not built")
    dev = weights[0].device
    ws = [w.detach().contiguous().float() for w in weights]
    bs = [b.detach().contiguous().float() for b in biases]
    for w, b in zip(ws, bs):
        _need_cuda(w, "weight")
        _need_cuda(b, "bias")
    packed = torch.empty(n, dtype=torch.float32, device=dev)
    wp = (ctypes.c_void_p * len(ws))(*[w.data_ptr() for w in ws])
    bp = (ctypes.c_void_p * len(bs))(*[b.data_ptr() for b in bs])
    check(L_.nsvf_mlp_train_pack(D, H, L, wp, bp, _p(packed), _stream(dev)), "mlp_train_pack")
    return packed


class FieldMLP(torch.autograd.Function):
    """The field MLP (field.py:217-279): L Linear + ReLU layers, then a
    4-wide head, output (sigma, sigmoid(r, g, b)).  Forward and backward are
    the hand-written FP32-MFMA kernels: the backward recomputes the
    activations from the input rather than storing them, and sums the
    weight gradients over samples in a fixed order (deterministic).
    Arguments: emb [M,D], then W_0, b_0, ..., W_L, b_L in torch layout."""

    @staticmethod
    def forward(ctx, emb, *params):
        weights, biases = list(params[0::2]), list(params[1::2])
        D, H, L = _mlp_shape(weights)
        _need_cuda(emb, "emb")
        emb = emb.contiguous()
        if emb.dtype != torch.float32 or emb.dim() != 2 or emb.shape[1] != D:
            raise RuntimeError(f"emb must be float32 [M, {D}]")
        packed = pack_mlp_train(weights, biases)
        M = emb.shape[0]
        out = torch.empty((M, 4), dtype=torch.float32, device=emb.device)
        check(lib().nsvf_mlp_train_forward(D, H, L, _p(packed), _p(emb), M, _p(out),
                                           _stream(emb.device)), "mlp_train_forward")
        ctx.save_for_backward(emb, packed)
        ctx.shape = (D, H, L)
        ctx.param_shapes = [tuple(t.shape) for t in params]
        return out

    @staticmethod
    def backward(ctx, dout):
        emb, packed = ctx.saved_tensors
        D, H, L = ctx.shape
        L_ = lib()
        dev = emb.device
        M = emb.shape[0]
        dout = dout.contiguous().float()
        demb = torch.empty((M, D), dtype=torch.float32, device=dev)
        n_part = int(L_.nsvf_mlp_train_partial_floats(D, H, L, M))
        n_par = int(L_.nsvf_mlp_train_param_floats(D, H, L))
        partial = torch.empty(n_part, dtype=torch.float32, device=dev)
        dparams = torch.empty(n_par, dtype=torch.float32, device=dev)
        check(L_.nsvf_mlp_train_backward(D, H, L, _p(packed), _p(emb), M, _p(dout), _p(demb),
                                         _p(partial), n_part, _p(dparams), _stream(dev)),
              "mlp_train_backward")
        grads, o = [], 0
        for shp in ctx.param_shapes:
            n = int(np.prod(shp))
            grads.append(dparams[o:o + n].view(shp))
            o += n
        return (demb, *grads)


_X3_INDEX = {}


def _x3_index(shapes, device):
    """For weights of the given [out, in] shapes: the position, in
    [0, W_0 flat, W_1 flat, ...] (0: zero padding), of every element of the
    split-BF16 kernel's fragments in their order, [fragments * 512] int64
    (pack_layer_bf16 applied to the positions; they stay below 2^24, so
    its float32 arrays carry them exactly)."""
    key = (tuple(shapes), str(device))
    idx = _X3_INDEX.get(key)
    if idx is None:
        parts, base = [], 1
        for O, K in shapes:
            pos = (np.arange(O * K, dtype=np.float64) + base).reshape(O, K).astype(np.float32)
            parts.append(pack_layer_bf16(pos, ((K + 31) // 32) * 32).reshape(-1))
            base += O * K
        if base >= 1 << 24:
            raise ValueError("split-BF16 training MLP: too many weights")
        idx = torch.from_numpy(np.concatenate(parts).astype(np.int64)).to(device)
        _X3_INDEX[key] = idx
    return idx


def pack_mlp_x3(weights, biases):
    """The split-BF16 MLP kernel's weight fragments and biases from
    torch-layout weights, on the device (the layout and rounding of
    field.pack_mlp's "bf16x3": per fragment its 512 hi parts bf16(w), then
    its lo parts bf16(w - hi), round to nearest even)."""
    dev = weights[0].device
    idx = _x3_index([tuple(w.shape) for w in weights], dev)
    flat = torch.cat([torch.zeros(1, dtype=torch.float32, device=dev)]
                     + [w.detach().reshape(-1).float() for w in weights])
    vals = flat[idx].view(-1, 512)
    hi = vals.to(torch.bfloat16)
    lo = (vals - hi.float()).to(torch.bfloat16)
    frags = torch.stack([hi, lo], 1).reshape(-1).view(torch.int16)
    head = biases[-1].detach().float()
    bias = torch.cat([b.detach().float() for b in biases[:-1]]
                     + [F.pad(head, (0, 16 - head.shape[0]))])
    return frags, bias


_X3T_INDEX = {}


def _x3_train_index(shapes, device):
    """For weights of the given [out, in] shapes (hidden layers, then the
    head): the position, in [0, W_0 flat, W_1 flat, ...] (0: zero padding),
    of every element of the split-BF16 backward's fragments in their order
    -- W_0 .. W_{L-1}, then their transposes -- [fragments * 512] int64."""
    key = (tuple(shapes), str(device))
    idx = _X3T_INDEX.get(key)
    if idx is None:
        pos, base = [], 1
        for O, K in shapes:
            pos.append((np.arange(O * K, dtype=np.float64) + base).reshape(O, K).astype(np.float32))
            base += O * K
        if base >= 1 << 24:
            raise ValueError("split-BF16 training MLP: too many weights")
        hidden = pos[:-1]
        parts = [pack_layer_bf16(q, ((q.shape[1] + 31) // 32) * 32).reshape(-1) for q in hidden]
        parts += [pack_layer_bf16(np.ascontiguousarray(q.T), ((q.shape[0] + 31) // 32) * 32)
                  .reshape(-1) for q in hidden]
        idx = torch.from_numpy(np.concatenate(parts).astype(np.int64)).to(device)
        _X3T_INDEX[key] = idx
    return idx


def pack_mlp_train_x3(weights):
    """The split-BF16 backward's weight fragments (nsvf_mlp_train_backward_x3):
    per fragment 512 hi parts bf16(w), then 512 lo parts bf16(w - hi), round
    to nearest even; an int16 device tensor."""
    dev = weights[0].device
    D, H, L = _mlp_shape(weights)
    idx = _x3_train_index([tuple(w.shape) for w in weights], dev)
    n = int(lib().nsvf_mlp_train_x3_frags(D, H, L))
    if n < 0 or idx.numel() != n * 512:
        raise RuntimeError(f"split-BF16 training backward: shape D={D} H={H} L={L} is not built")
    flat = torch.cat([torch.zeros(1, dtype=torch.float32, device=dev)]
                     + [w.detach().reshape(-1).float() for w in weights])
    vals = flat[idx].view(-1, 512)
    hi = vals.to(torch.bfloat16)
    lo = (vals - hi.float()).to(torch.bfloat16)
    return torch.stack([hi, lo], 1).reshape(-1).view(torch.int16)


class FieldMLPx3(FieldMLP):
    """FieldMLP with its forward on the renderer's split-BF16 MLP kernel
    (three BF16 MFMAs per product on hi / lo operand parts, FP32
    accumulation: outputs within about 1e-5 of FP32) -- the gathered
    features written as its operand tiles, the weights packed on the device
    each call -- and the backward kernel with the activations'
    recomputation, the input gradients and the parameter gradients in split
    BF16 too (its head's recomputed logits and input gradient in FP32),
    deterministic like FieldMLP's."""

    @staticmethod
    def forward(ctx, emb, *params):
        weights, biases = list(params[0::2]), list(params[1::2])
        D, H, L = _mlp_shape(weights)
        _need_cuda(emb, "emb")
        emb = emb.contiguous()
        if emb.dtype != torch.float32 or emb.dim() != 2 or emb.shape[1] != D:
            raise RuntimeError(f"emb must be float32 [M, {D}]")
        L_ = lib()
        if not L_.nsvf_field_bf16x3_supported(D, H, L):
            raise RuntimeError(f"split-BF16 MLP not built for D={D} H={H} L={L}")
        dev, M = emb.device, emb.shape[0]
        out = torch.empty((M, 4), dtype=torch.float32, device=dev)
        if M > 0:
            frags, bias = pack_mlp_x3(weights, biases)
            ws_bytes = int(L_.nsvf_field_workspace_bytes(2, D, H, L, M))
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
            m_dev = torch.full((1,), M, dtype=torch.int32, device=dev)
            st = _stream(dev)
            check(L_.nsvf_field_tiles_bf16x3(D, H, L, _p(emb), M, _p(ws), ws_bytes, st),
                  "field_tiles_bf16x3")
            check(L_.nsvf_field_mlp_bf16x3(D, H, L, _p(ws), ws_bytes, _p(m_dev), M, _p(frags),
                                           _p(bias), _p(out), None, st), "field_mlp_bf16x3")
        ctx.save_for_backward(emb, pack_mlp_train(weights, biases), pack_mlp_train_x3(weights))
        ctx.shape = (D, H, L)
        ctx.param_shapes = [tuple(t.shape) for t in params]
        return out

    @staticmethod
    def backward(ctx, dout):
        emb, packed, x3w = ctx.saved_tensors
        D, H, L = ctx.shape
        L_ = lib()
        dev = emb.device
        M = emb.shape[0]
        dout = dout.contiguous().float()
        demb = torch.empty((M, D), dtype=torch.float32, device=dev)
        n_part = int(L_.nsvf_mlp_train_partial_floats(D, H, L, M))
        n_par = int(L_.nsvf_mlp_train_param_floats(D, H, L))
        partial = torch.empty(n_part, dtype=torch.float32, device=dev)
        dparams = torch.empty(n_par, dtype=torch.float32, device=dev)
        check(L_.nsvf_mlp_train_backward_x3(D, H, L, _p(packed), _p(x3w), _p(emb), M, _p(dout),
                                            _p(demb), _p(partial), n_part, _p(dparams),
                                            _stream(dev)), "mlp_train_backward_x3")
        grads, o = [], 0
        for shp in ctx.param_shapes:
            n = int(np.prod(shp))
            grads.append(dparams[o:o + n].view(shp))
            o += n
        return (demb, *grads)


class TrainableField(nn.Module):
    """The trainable state of a voxel scene: corner embeddings [K,D] and the
    field MLP (L hidden ReLU layers, a 4-wide head: sigma, sigmoid RGB).

    mlp="hip": the MLP runs forward and backward on the hand-written
    FP32-MFMA kernels (FieldMLP; built for D=32, H=128, L=3 -- other shapes
    raise); mlp="hip_x3": forward and backward in split BF16 (FieldMLPx3:
    outputs within about 1e-5 of FP32, gradients about as close; the
    backward's head -- recomputed logits and their input gradient -- stays
    FP32);
    mlp="torch": torch.nn.functional.linear with torch autograd, as in the
    reference."""

    def __init__(self, scene: VoxelScene, device, mlp="hip"):
        super().__init__()
        dev = torch.device(device)
        if mlp not in ("hip", "hip_x3", "torch"):
            raise ValueError(f"mlp={mlp!r}: 'hip', 'hip_x3' or 'torch'")
        self.mlp_impl = mlp
        self.values = nn.Parameter(torch.from_numpy(np.array(scene.values, np.float32)).to(dev))
        self.weights = nn.ParameterList(
            [nn.Parameter(torch.from_numpy(np.array(w, np.float32)).to(dev))
             for w in scene.mlp.weights])
        self.biases = nn.ParameterList(
            [nn.Parameter(torch.from_numpy(np.array(b, np.float32)).to(dev))
             for b in scene.mlp.biases])

    def mlp(self, emb):
        """[M,D] -> [M,4] = (sigma, r, g, b)."""
        if self.mlp_impl in ("hip", "hip_x3"):
            D, H, L = _mlp_shape(list(self.weights))
            if not mlp_train_supported(D, H, L):
                raise RuntimeError(f"the HIP training MLP is not built for D={D} H={H} L={L}; "
                                   "construct the field with mlp='torch'")
            params = []
            for w, b in zip(self.weights, self.biases):
                params += [w, b]
            fn = FieldMLPx3 if self.mlp_impl == "hip_x3" else FieldMLP
            return fn.apply(emb, *params)
        h = emb
        n = len(self.weights)
        for i in range(n):
            h = F.linear(h, self.weights[i], self.biases[i])
            if i < n - 1:
                h = torch.relu(h)
        return torch.cat([h[:, :1], torch.sigmoid(h[:, 1:4])], 1)


class TrainRenderer:
    """Differentiable view rendering against a DeviceScene's geometry (its
    octree, shifted centres and corner ids) with a TrainableField's
    parameters: intersect + coarse samples (non-differentiable kernels), then
    gather, MLP and compositing with autograd, then the background fill
    (nsvf.py:94-101).  coherent: the traversal takes the batch's rays
    grouped by direction and origin (coherent_order; a batch of random
    pixels from several views has no image rows to tile), which changes no
    result."""

    def __init__(self, device, compat=True, deterministic_step=False, seed=0, coherent=True):
        self.device = torch.device(device)
        self.geometry = Renderer(device, precision="fp32", hierarchical=False, compat=compat)
        self.coherent = bool(coherent)
        # training samples with noise unless --deterministic-step
        # (encoder.py:546: deterministic = deterministic_step or not training);
        # forward() k of this renderer uses noise seed `seed + k` (mod 2^32)
        self.deterministic_step = bool(deterministic_step)
        self.seed = int(seed)
        self.steps = 0
        self.last = None
        self.last_seed = None

    def forward(self, sc: DeviceScene, field: TrainableField, ray_start, ray_dir, seed=None):
        """ray_start, ray_dir [R,3] float32 on the device.  Returns (rgb [R,3],
        depth [R], missed [R]) connected to `field`'s parameters.  seed: this
        step's noise seed (default: the renderer's next one); ignored with
        deterministic_step."""
        _need_cuda(ray_start, "ray_start")
        _need_cuda(ray_dir, "ray_dir")
        rs, rd = ray_start.contiguous().float(), ray_dir.contiguous().float()
        if self.deterministic_step:
            seed = None
        elif seed is None:
            seed = (self.seed + self.steps) & 0xFFFFFFFF
        self.steps += 1
        self.last_seed = seed
        self.geometry.noise_seed = seed
        order = coherent_order(rs, rd) if self.coherent and rs.shape[0] > 0 else None
        g = self.geometry.coarse_stream(sc, rs, rd, order=order).owned()
        self.last = g
        rec = g.samples
        s_idx = rec[:, 0]
        depth = rec[:, 1].view(torch.float32)
        ray = rec[:, 3].long()
        # sample points x = o + d * z, each operation rounded (geometry.py:160-161)
        xyz = rs.index_select(0, ray) + rd.index_select(0, ray) * depth[:, None]
        if sc.gather_plan is None:  # the scene's incidence order, once per scene
            sc.gather_plan = gather_backward_plan(sc.feats, field.values.shape[0])
        emb = trilinear_gather(field.values, s_idx, xyz, sc.centers, sc.feats, sc.voxel_size,
                               sc.gather_plan)
        out = field.mlp(emb)
        rgb, dep, missed = StreamComposite.apply(out, g)
        bg = torch.tensor(BG_COLOR, dtype=torch.float32, device=rgb.device)
        return rgb + missed[:, None] * bg, dep + missed * BG_DEPTH, missed


def mse_loss(rgb, target):
    """Mean squared RGB error over rays and channels (rendering_loss.py:139
    with the RGB term only)."""
    return torch.mean((rgb - target) ** 2)
