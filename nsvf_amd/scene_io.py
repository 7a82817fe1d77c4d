"""Scene files of the reference, read and written on the host:

* the initial grid of a bounding-box file (``bbox.txt``: ``x_min y_min z_min
  x_max y_max z_max voxel_size``; fairnr/modules/encoder.py:263-267,
  ``bbox2voxels`` encoder.py:1053-1059);
* voxel ``.ply`` files -- vertices (+ the ``quality`` voxel index the
  exporter writes) or a voxel mesh, whose first face gives the voxel size
  (encoder.py:218-258) -- and old-style ``.txt`` point lists (encoder.py:
  260-261); the exporter of both ``.ply`` forms (encoder.py:412-454);
* the encoder's checkpoint buffers (encoder.py:285-293, ``values.weight``
  of its ``Embedding``), read and written under the model's ``encoder.``
  prefix; pruned voxels (``keep`` = 0) are dropped on load, as
  ``precompute`` does (encoder.py:378-383).

The voxel corner keys are built as the encoder's constructor does
(encoder.py:269-276).  PLY is read and written by the small codec below
(ASCII and binary, scalar and list properties); the field MLP is not part of
these files (a ``FieldMLP`` is passed in, or made from a seed).
"""
from __future__ import annotations

import os

import numpy as np

from .scene import FieldMLP, VoxelScene, _corner_keys, make_mlp

# ---------------------------------------------------------------------------
# PLY codec

_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2",
              "int16": "i2", "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4",
              "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8",
              "float64": "f8"}
_PLY_NAMES = {"i1": "char", "u1": "uchar", "i2": "short", "u2": "ushort", "i4": "int",
              "u4": "uint", "f4": "float", "f8": "double"}


def read_ply(path):
    """{element name: {property name: array}} of a PLY file; list
    properties come back as a list of arrays (or a 2-D array when all rows
    have the same length)."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt, elements = None, []
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: header without end_header")
            tok = line.decode("ascii").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                elements.append((tok[1], int(tok[2]), []))
            elif tok[0] == "property":
                if tok[1] == "list":
                    elements[-1][2].append((tok[4], _PLY_TYPES[tok[2]], _PLY_TYPES[tok[3]]))
                else:
                    elements[-1][2].append((tok[2], _PLY_TYPES[tok[1]], None))
            elif tok[0] == "end_header":
                break
        if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
            raise ValueError(f"{path}: unsupported PLY format {fmt}")
        end = "<" if fmt == "binary_little_endian" else ">"
        out = {}
        if fmt == "ascii":
            words = f.read().decode("ascii").split()
            pos = 0
            for name, n, props in elements:
                cols = {p[0]: [] for p in props}
                for _ in range(n):
                    for pname, t, lt in props:
                        if lt is None:
                            cols[pname].append(words[pos])
                            pos += 1
                        else:
                            k = int(words[pos])
                            cols[pname].append(np.array(words[pos + 1:pos + 1 + k], dtype=lt))
                            pos += 1 + k
                out[name] = {p[0]: (np.array(cols[p[0]], dtype=p[1]) if p[2] is None
                                    else _lists(cols[p[0]])) for p in props}
            return out
        data = f.read()
    pos = 0
    for name, n, props in elements:
        if all(p[2] is None for p in props):
            dt = np.dtype([(p[0], end + p[1]) for p in props])
            arr = np.frombuffer(data, dtype=dt, count=n, offset=pos)
            pos += n * dt.itemsize
            out[name] = {p[0]: arr[p[0]].astype(p[1]) for p in props}
            continue
        cols = {p[0]: [] for p in props}
        for _ in range(n):
            for pname, t, lt in props:
                if lt is None:
                    v = np.frombuffer(data, end + t, 1, pos)[0]
                    pos += np.dtype(t).itemsize
                    cols[pname].append(v)
                else:
                    k = int(np.frombuffer(data, end + t, 1, pos)[0])
                    pos += np.dtype(t).itemsize
                    cols[pname].append(np.frombuffer(data, end + lt, k, pos).astype(lt))
                    pos += k * np.dtype(lt).itemsize
        out[name] = {p[0]: (np.array(cols[p[0]], dtype=p[1]) if p[2] is None
                            else _lists(cols[p[0]])) for p in props}
    return out


def _lists(rows):
    if rows and all(len(r) == len(rows[0]) for r in rows):
        return np.stack(rows)
    return rows


def write_ply(path, elements, text=False):
    """Writes {element: [(property, array)]} (arrays of length n; a 2-D
    integer array is written as a list property with uchar counts), binary
    little-endian unless text."""
    lines = ["ply", "format " + ("ascii" if text else "binary_little_endian") + " 1.0"]
    for name, props in elements.items():
        n = len(props[0][1])
        lines.append(f"element {name} {n}")
        for pname, arr in props:
            arr = np.asarray(arr)
            t = _PLY_NAMES[arr.dtype.str[1:]]
            lines.append(f"property list uchar {t} {pname}" if arr.ndim == 2 else
                         f"property {t} {pname}")
    lines.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode("ascii"))
        for name, props in elements.items():
            n = len(props[0][1])
            arrs = [np.asarray(a) for _, a in props]
            if text:
                for i in range(n):
                    row = []
                    for a in arrs:
                        row += ([str(a.shape[1])] + [repr(x.item()) for x in a[i]] if a.ndim == 2
                                else [repr(a[i].item())])
                    f.write((" ".join(row) + "\n").encode("ascii"))
                continue
            fields = []
            for (pname, _), a in zip(props, arrs):
                if a.ndim == 2:
                    fields += [(pname + "__n", "u1"), (pname, "<" + a.dtype.str[1:], (a.shape[1],))]
                else:
                    fields.append((pname, "<" + a.dtype.str[1:]))
            rec = np.zeros(n, dtype=fields)
            for (pname, _), a in zip(props, arrs):
                if a.ndim == 2:
                    rec[pname + "__n"] = a.shape[1]
                rec[pname] = a
            f.write(rec.tobytes())


# ---------------------------------------------------------------------------
# voxel sources

def bbox2voxels(bbox, voxel_size):
    """encoder.py:1053-1059: voxel centres on the grid from bbox[:3] with
    spacing voxel_size up to bbox[3:6] (inclusive, rounded), in
    ``np.meshgrid``'s default (x-y swapped) order, float32."""
    bbox = np.asarray(bbox, np.float64)
    vox_min, vox_max = bbox[:3], bbox[3:6]
    steps = np.round((vox_max - vox_min) / voxel_size).astype(np.int64) + 1
    x, y, z = [c.reshape(-1).astype(np.float32)
               for c in np.meshgrid(np.arange(steps[0]), np.arange(steps[1]), np.arange(steps[2]))]
    # float32 arithmetic: a float32 array with a float64 scalar stays float32
    # under the NumPy the reference was written for
    vs, m = np.float32(voxel_size), vox_min.astype(np.float32)
    x, y, z = x * vs + m[0], y * vs + m[1], z * vs + m[2]
    return np.stack([x, y, z]).T.astype(np.float32)


def read_bbox(path):
    """``bbox.txt``: (bbox [6], voxel size from the 7th number)."""
    b = np.loadtxt(path, dtype=np.float64).reshape(-1)
    if b.size < 7:
        raise ValueError(f"{path}: expected x_min y_min z_min x_max y_max z_max voxel_size")
    return b[:6], float(b[6])


def load_voxels(path, voxel_size=None, kind=None):
    """Voxel centres of a ``.ply`` / ``.txt`` voxel file or a bounding-box
    file (encoder.py:215-267); kind "ply", "txt" or "bbox" (default: by
    extension, ``bbox.txt`` for the box).  Returns (points [N,3] f32, voxel
    size, voxel index or None) -- the index is the ``quality`` property the
    exporter writes, which a caller may use to reorder checkpoint buffers
    (encoder.py:320-323).  The reference tests ``'quality'`` among the file's
    element names (encoder.py:257), where it never is, so it never reorders;
    the index here is an explicit opt-in.  A mesh-form file gives no index:
    its points are the unique corner vertices, not one row per voxel."""
    if kind is None:
        kind = ("ply" if path.endswith(".ply") else
                "bbox" if os.path.basename(path) == "bbox.txt" else "txt")
    if kind == "ply":
        ply = read_ply(path)
        if "vertex" not in ply:
            raise ValueError(f"{path}: no vertex element")
        v = ply["vertex"]
        pts = np.stack([v["x"], v["y"], v["z"]]).astype(np.float32).T
        index = None
        if "face" in ply:
            faces = ply["face"]["vertex_indices"]
            t = pts[np.asarray(faces[0]).astype(np.int64)]
            voxel_size = float(np.abs(t[0] - t[1]).max())
            pts = np.unique(pts, axis=0)
        elif voxel_size is None:
            raise ValueError(f"{path}: a vertex-only voxel file needs the voxel size")
        elif "quality" in v:
            index = v["quality"].astype(np.int64)
        return np.ascontiguousarray(pts), float(voxel_size), index
    if kind == "bbox":
        bbox, vs = read_bbox(path)
        vs = vs if voxel_size is None else float(voxel_size)
        return bbox2voxels(bbox, vs), vs, None
    if kind != "txt":
        raise ValueError(f"unknown voxel file kind {kind}")
    if voxel_size is None:
        raise ValueError(f"{path}: a .txt voxel file needs the voxel size")
    pts = np.loadtxt(path)[:, 3:].astype(np.float32)
    return np.ascontiguousarray(pts), float(voxel_size), None


def scene_from_points(points, voxel_size, embed_dim=32, step_ratio=0.125, step_size=None,
                      max_hits=60, mlp: FieldMLP = None, values=None, seed=0, name=""):
    """A VoxelScene over voxel centres, with the encoder's corner keys
    (encoder.py:269-276) and, unless given, values drawn as its Embedding
    does (normal, std embed_dim ** -0.5; module_utils.py:23-26) and a
    3 x 128 ReLU field."""
    points = np.ascontiguousarray(points, np.float32)
    keys, feats = _corner_keys(points, voxel_size)
    if values is None:
        rng = np.random.default_rng(seed)
        values = (rng.standard_normal((keys.shape[0], embed_dim)) * embed_dim ** -0.5).astype(np.float32)
    values = np.ascontiguousarray(values, np.float32)
    if mlp is None:
        mlp = make_mlp(values.shape[1], 128, 3, seed + 1)
    vs = float(np.float32(voxel_size))
    ss = float(np.float32(step_size if step_size is not None else vs * step_ratio))
    return VoxelScene(points=points, voxel_size=vs, step_size=ss, max_hits=int(max_hits),
                      embed_dim=values.shape[1], keys=keys, feats=feats, values=values, mlp=mlp,
                      name=name)


def export_voxels(scene: VoxelScene, path, keep=None, return_mesh=False):
    """encoder.py:412-454: the kept voxel centres with their original index
    as ``quality``, or (return_mesh) the voxels as a quad mesh -- corner
    vertices in first-seen order, one face per distinct face centre."""
    n = scene.num_voxels
    keep = np.ones(n, bool) if keep is None else np.asarray(keep, bool)
    idx = np.arange(n)[keep]
    pts = scene.points[keep]
    if not return_mesh:
        write_ply(path, {"vertex": [("x", pts[:, 0].astype(np.float32)),
                                    ("y", pts[:, 1].astype(np.float32)),
                                    ("z", pts[:, 2].astype(np.float32)),
                                    ("quality", idx.astype(np.float32))]})
        return
    vs = np.float32(scene.voxel_size)
    half = vs / np.float32(2)
    mn = pts.min(0, keepdims=True)
    coords = np.round((pts - mn) / half).astype(np.int64)
    residual = (pts - coords.astype(np.float32) * half).mean(0, keepdims=True, dtype=np.float32)
    offsets = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1], [1, 1, -1],
                        [1, -1, 1], [-1, 1, 1], [1, 1, 1]], np.int64)
    vcoords = coords[:, None, :] + offsets[None]                       # [n,8,3]
    vpoints = (vcoords.astype(np.float32) * vs) / np.float32(2) + residual
    flat = vcoords.reshape(-1, 3)
    _, first, inv = np.unique(flat, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                                          # first-seen order
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    vid = rank[inv.reshape(-1)].reshape(-1, 8)
    verts = vpoints.reshape(-1, 3)[first[order]]
    faceidxs = np.array([[1, 6, 7, 5], [7, 6, 2, 4], [5, 7, 4, 3], [1, 0, 2, 6], [1, 5, 3, 0],
                         [0, 3, 4, 2]])
    fcorner = vcoords[:, faceidxs]                                     # [n,6,4,3]
    fkey = fcorner.sum(2) // 4                                         # face centres
    fverts = vid[:, faceidxs].reshape(-1, 4)
    _, ffirst = np.unique(fkey.reshape(-1, 3), axis=0, return_index=True)
    faces = fverts[np.sort(ffirst)].astype(np.int32)
    write_ply(path, {"vertex": [("x", verts[:, 0]), ("y", verts[:, 1]), ("z", verts[:, 2])],
                     "face": [("vertex_indices", faces)]})


# ---------------------------------------------------------------------------
# checkpoints

def state_dict(scene: VoxelScene, keep=None, prefix="encoder."):
    """The encoder's buffers and embedding (encoder.py:285-293, 310): names,
    dtypes and shapes as the reference registers them."""
    import torch
    n = scene.num_voxels
    keep = np.ones(n, np.int64) if keep is None else np.asarray(keep, np.int64)
    return {
        prefix + "points": torch.from_numpy(np.ascontiguousarray(scene.points, np.float32)),
        prefix + "keys": torch.from_numpy(np.ascontiguousarray(scene.keys, np.int64)),
        prefix + "feats": torch.from_numpy(np.ascontiguousarray(scene.feats, np.int64)),
        prefix + "num_keys": torch.tensor(int(scene.keys.shape[0]), dtype=torch.int64),
        prefix + "keep": torch.from_numpy(keep),
        prefix + "voxel_size": torch.tensor(scene.voxel_size, dtype=torch.float32),
        prefix + "step_size": torch.tensor(scene.step_size, dtype=torch.float32),
        prefix + "max_hits": torch.tensor(float(scene.max_hits), dtype=torch.float32),
        prefix + "values.weight": torch.from_numpy(np.ascontiguousarray(scene.values, np.float32)),
    }


def save_checkpoint(scene: VoxelScene, path, keep=None, prefix="encoder."):
    """A checkpoint file holding the model state under "model", as a fairseq
    checkpoint does."""
    import torch
    torch.save({"model": state_dict(scene, keep, prefix)}, path)


def load_checkpoint(src, prefix="encoder.", mlp: FieldMLP = None, voxel_index=None,
                    name=""):
    """A VoxelScene from a checkpoint path (loaded with weights_only=True)
    or a state dict: kept voxels only (encoder.py:378-383), values cut to
    num_keys; voxel_index (a voxel file's ``quality``) reorders the voxel
    buffers first (encoder.py:320-323).  The field MLP is not restated
    from the checkpoint: pass ``mlp`` (default: a seeded 3 x 128 field)."""
    import torch
    sd = src
    if not isinstance(src, dict):
        sd = torch.load(src, map_location="cpu", weights_only=True)
    if "model" in sd and not any(k.startswith(prefix) for k in sd):
        sd = sd["model"]

    def get(k):
        if prefix + k not in sd:
            raise KeyError(f"checkpoint has no {prefix + k}")
        t = sd[prefix + k]
        return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
    points, feats, keep = get("points"), get("feats"), get("keep")
    if voxel_index is not None:
        points, feats, keep = points[voxel_index], feats[voxel_index], keep[voxel_index]
    kb = keep.astype(bool)
    nk = int(get("num_keys"))
    values = get("values.weight")[:nk].astype(np.float32)
    if mlp is None:
        mlp = make_mlp(values.shape[1], 128, 3, 1)
    return VoxelScene(points=np.ascontiguousarray(points[kb], np.float32),
                      voxel_size=float(get("voxel_size")), step_size=float(get("step_size")),
                      max_hits=int(float(get("max_hits"))), embed_dim=values.shape[1],
                      keys=np.ascontiguousarray(get("keys"), np.int64),
                      feats=np.ascontiguousarray(feats[kb], np.int64),
                      values=np.ascontiguousarray(values), mlp=mlp, name=name)
