"""ctypes binding of ``libnsvf_amd.so`` (the C ABI declared in
``include/nsvf_amd.h``).

The library is built in-tree by ``python -m nsvf_amd.build``.  There is no
fallback: if the library is missing or lacks a symbol, importing an op raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsvf_amd.so")

_lock = threading.Lock()
_LIB = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_L = ctypes.c_int64


class Camera(ctypes.Structure):
    """``nsvf_camera``: fx, fy, cx, cy; inv_rt (3 x 4 row-major); H, W, h, w."""
    _fields_ = [("fx", _F), ("fy", _F), ("cx", _F), ("cy", _F), ("inv_rt", _F * 12),
                ("H", _I), ("W", _I), ("h", _I), ("w", _I)]


_CAM = ctypes.POINTER(Camera)

# name -> (argtypes, restype); every entry point declared in include/nsvf_amd.h
SIGNATURES = {
    "nsvf_last_error": ([], ctypes.c_char_p),
    "nsvf_version": ([], _I),
    "nsvf_aabb_intersect": ([_P, _P, _P, _I, _I, _I, _I, _F, _I, _P, _P, _P, _P], _I),
    "nsvf_svo_intersect": ([_P, _P, _P, _P, _I, _I, _I, _I, _F, _I, _P, _P, _P, _P], _I),
    "nsvf_build_octree": ([_P, _P, _I, _I, ctypes.POINTER(_P), ctypes.POINTER(_I)], _I),
    "nsvf_octree_take": ([_P, _P, _P], _I),
    "nsvf_inverse_cdf_sampling": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _I, _P, _P, _P, _P, _P], _I),
    "nsvf_inverse_cdf_sampling_rows": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _F, _I, _P, _P, _P, _P, _P], _I),
    "nsvf_uniform_ray_sampling": ([_P, _P, _P, _P, _I, _I, _I, _I, _F, _P, _P, _P, _P, _P], _I),
    "nsvf_trilinear_gather": ([_P, _P, _L, _P, _P, _P, _I, _F, _P, _P], _I),
    "nsvf_trilinear_gather_backward_workspace": ([_L, _L, _L, _I, _P], _L),
    "nsvf_trilinear_gather_backward": ([_P, _P, _L, _P, _P, _L, _L, _I, _F, _P, _P, _P, _L, _P],
                                       _I),
    "nsvf_trilinear_gather_backward_plan_bytes": ([_L, _L], _L),
    "nsvf_trilinear_gather_backward_plan": ([_P, _L, _L, _P, _L, _P, _L, _P], _I),
    "nsvf_trilinear_gather_backward_planned": ([_P, _P, _L, _P, _P, _L, _L, _I, _F, _P, _P, _P,
                                                _P, _L, _P], _I),
    "nsvf_field_prune": ([_I, _I, _I, _I, _P, _P, _P, _L, _F, _L, _L, _I, _P, _P, _P, _P], _I),
    "nsvf_split_workspace_bytes": ([_L, _P], _L),
    "nsvf_split_voxels": ([_P, _L, _F, _P, _L, _P, _P, _P, _P], _I),
    "nsvf_split_values": ([_P, _P, _P, _L, _I, _F, _P, _L, _L, _P, _P, _P], _I),
    "nsvf_scan_workspace_ints": ([_L], _L),
    "nsvf_scan_exclusive_i32": ([_P, _P, _L, _P, _P, _P], _I),
    "nsvf_render_hits": ([_P, _P, _L, _I, _P, _I, _I, _I, _F, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_render_hits_ordered": ([_P, _P, _L, _P, _P, _I, _I, _I, _F, _P, _P, _P, _P, _P, _P],
                                 _I),
    "nsvf_ray_order_keys": ([_P, _P, _L, _P, _P, _P], _I),
    "nsvf_generate_rays": ([_CAM, _P, _P, _P], _I),
    "nsvf_render_hits_camera": ([_CAM, _P, _I, _I, _I, _F, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_render_compact": ([_P, _P, _L, _P, _P], _I),
    "nsvf_render_sample_coarse": ([_P, _P, _L, _P, _P, _P, _P, _P, _I, _F, _I, _F, _P, _P, _P,
                                   _P, _P, _I, _P, _L, _P], _I),
    "nsvf_render_sample_fine": ([_P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, _F, _I, _P,
                                 _P, _P, _P, _I, _P, _L, _P], _I),
    "nsvf_render_eval_list_blocks": ([_L], _L),
    "nsvf_march_hist_ints": ([_L, _I, _I], _L),
    "nsvf_march_begin": ([_P, _L, _P, _I, _I, _P, _P, _P, _P], _I),
    "nsvf_march_step": ([_P, _L, _P, _P, _P, _P, _I, _I, _F, _P, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_march_end": ([_P, _L, _P, _P, _P, _P, _P], _I),
    "nsvf_mlp_train_supported": ([_I, _I, _I], _I),
    "nsvf_mlp_train_packed_floats": ([_I, _I, _I], _L),
    "nsvf_mlp_train_param_floats": ([_I, _I, _I], _L),
    "nsvf_mlp_train_partial_floats": ([_I, _I, _I, _L], _L),
    "nsvf_mlp_train_pack": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "nsvf_mlp_train_forward": ([_I, _I, _I, _P, _P, _L, _P, _P], _I),
    "nsvf_mlp_train_backward": ([_I, _I, _I, _P, _P, _L, _P, _P, _P, _L, _P, _P], _I),
    "nsvf_mlp_train_x3_frags": ([_I, _I, _I], _L),
    "nsvf_mlp_train_backward_x3": ([_I, _I, _I, _P, _P, _P, _L, _P, _P, _P, _L, _P, _P], _I),
    "nsvf_render_eval_list_count": ([_P, _P, _L, _P, _P], _I),
    "nsvf_render_eval_list": ([_P, _P, _L, _P, _P, _P], _I),
    "nsvf_render_reuse_fill": ([_P, _P, _L, _P, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_render_composite_reuse": ([_P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     _P], _I),
    "nsvf_field_supported": ([_I, _I, _I], _I),
    "nsvf_field_fused_supported": ([_I, _I, _I], _I),
    "nsvf_field_workspace_bytes": ([_I, _I, _I, _I, _L], _L),
    "nsvf_field_eval": ([_I, _I, _I, _I, _P, _P, _P, _F, _P, _P, _P, _P, _P, _L, _P, _P, _P, _P,
                         _L, _P], _I),
    "nsvf_field_embed_bf16": ([_I, _I, _I, _P, _P, _P, _F, _P, _P, _P, _P, _P, _L, _P, _L, _P],
                              _I),
    "nsvf_field_mlp_bf16": ([_I, _I, _I, _P, _L, _P, _L, _P, _P, _P, _P, _P], _I),
    "nsvf_field_bf16x3_supported": ([_I, _I, _I], _I),
    "nsvf_field_static_scratch_bytes": ([], _L),
    "nsvf_field_eval_static": ([_I, _I, _I, _I, _P, _P, _P, _F, _P, _P, _P, _P, _P, _L, _P, _P,
                                _P, _P], _I),
    "nsvf_field_embed_bf16x3": ([_I, _I, _I, _P, _P, _P, _F, _P, _P, _P, _P, _P, _L, _P, _L,
                                 _P], _I),
    "nsvf_field_mlp_bf16x3": ([_I, _I, _I, _P, _L, _P, _L, _P, _P, _P, _P, _P], _I),
    "nsvf_field_tiles_bf16x3": ([_I, _I, _I, _P, _L, _P, _L, _P], _I),
    "nsvf_render_composite": ([_P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_render_composite_backward": ([_P, _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "nsvf_track_voxel_probs": ([_P, _P, _L, _P, _P, _P, _P, _P, _L, _P], _I),
    "nsvf_render_finish": ([_L, _P, _P, _P, _P, _F, _F, _F, _F, _P, _P, _P, _P], _I),
    "nsvf_plain_aabb_intersect": ([_P, _P, _P, _I, _I, _I, _F, _I, _P, _P, _P, _P], _I),
    "nsvf_plain_svo_intersect": ([_P, _P, _P, _P, _I, _I, _I, _F, _I, _P, _P, _P, _P], _I),
    "nsvf_plain_inverse_cdf_sampling": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _I, _P, _P,
                                         _P, _P], _I),
}


class NSVFError(RuntimeError):
    """Error returned by the native library (bad argument or HIP failure)."""


def lib():
    """Load the shared library once; raise if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    with _lock:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not built; run `python -m nsvf_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (argtypes, restype) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _LIB = L
        return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().nsvf_last_error()
        raise NSVFError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
