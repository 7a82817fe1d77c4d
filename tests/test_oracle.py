"""Pins for the CPU oracle: closed forms, hand-computed cases, invariants and
brute force.  The reference ships no tests or golden vectors, so these are
what ties the oracle to the reference's algorithm; each targets a plausible
transcription mistake (wrong octant bit order, BFS numbering, missing tail
sample, wrong trailing-loop stop, swapped slab bounds, dropped x7, ...)."""
import math

import numpy as np
import pytest

import nsvf_oracle as O
from _cases import axis_rays, random_rays, shell, shell_octree, shell_rays
from nsvf_amd import scene as S


# ---------------------------------------------------------------------------
# Slab test / intersection

def dyadic_grid(res=8):
    vs = np.float32(2.0 / res)
    c = (np.float32(-1) + vs * (np.arange(res, dtype=np.float32) + np.float32(0.5)))
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32), vs


def test_axis_ray_closed_form():
    pts, vs = dyadic_grid(8)
    keep = np.random.default_rng(0).random(len(pts)) < 0.6
    pts = pts[keep]
    y0, z0 = np.float32(-0.375 + 0.05), np.float32(0.125 + 0.03)
    rs = np.array([[[-3.0, y0, z0]]], np.float32)
    rd = np.array([[[1.0, 0.0, 0.0]]], np.float32)
    idx, lo, hi = O.aabb_intersect(rs, rd, pts[None], vs, 60)
    h = vs / 2
    row = np.nonzero((np.abs(pts[:, 1] - y0) < h) & (np.abs(pts[:, 2] - z0) < h))[0]
    n = len(row)
    assert n > 0
    assert list(idx[0, 0, :n]) == list(row)            # voxel-index order
    assert (idx[0, 0, n:] == -1).all()
    assert np.array_equal(lo[0, 0, :n], pts[row, 0] - h + np.float32(3))   # exact (dyadic)
    assert np.array_equal(hi[0, 0, :n], pts[row, 0] + h + np.float32(3))
    assert (lo[0, 0, n:] == 0).all() and (hi[0, 0, n:] == 0).all()


def test_ray_on_face_plane_hits_both_rows():
    """Direction 0 on an axis and origin exactly on a face plane: that axis
    gives NaN bounds and is skipped, so both adjacent rows are hit."""
    pts, vs = dyadic_grid(8)
    y_plane = np.float32(0.0)   # the face between rows 3 and 4
    rs = np.array([[[-3.0, y_plane, 0.1]]], np.float32)
    rd = np.array([[[1.0, 0.0, 0.0]]], np.float32)
    idx, lo, hi = O.aabb_intersect(rs, rd, pts[None], vs, 60)
    h = vs / 2
    rows = (np.abs(np.abs(pts[:, 1]) - h) < 1e-6) & (np.abs(pts[:, 2] - 0.125) < 1e-6)
    assert (idx[0, 0] != -1).sum() == rows.sum() == 16


def test_entry_clipped_at_origin_inside_box():
    pts = np.array([[[0.0, 0.0, 0.0]]], np.float32)
    rs = np.array([[[0.01, 0.02, 0.03]]], np.float32)
    rd = np.array([[[0.0, 0.0, 1.0]]], np.float32)
    idx, lo, hi = O.aabb_intersect(rs, rd, pts, 1.0, 4)
    assert idx[0, 0, 0] == 0 and lo[0, 0, 0] == 0.0
    assert hi[0, 0, 0] == np.float32(0.5) - np.float32(0.03)


def test_fp32_hits_agree_with_fp64_except_grazing():
    sc = shell()
    rs, rd = shell_rays(width=32)
    pts = O.shift_centers(sc.points, sc.voxel_size)
    i32, lo32, hi32 = O.aabb_intersect(rs[None], rd[None], pts[None], sc.voxel_size, 60)
    i64, lo64, hi64 = O.aabb_intersect_f64(rs, rd, pts, sc.voxel_size, 60)
    diff_rays = [r for r in range(rs.shape[0]) if set(i32[0, r]) != set(i64[r])]
    assert len(diff_rays) <= 3
    for r in diff_rays:  # a grazing ray: the differing voxel is entered over < 1e-5
        only = (set(i32[0, r]) ^ set(i64[r])) - {-1}
        for k in only:
            j = list(i64[r]).index(k) if k in i64[r] else None
            if j is not None:
                assert hi64[r, j] - lo64[r, j] < 1e-5
    same = [r for r in range(rs.shape[0]) if r not in diff_rays]
    m = i32[0, same] != -1
    assert np.abs(lo32[0, same][m] - lo64[same][m]).max() < 1e-5
    assert np.abs(hi32[0, same][m] - hi64[same][m]).max() < 1e-5


@pytest.mark.parametrize("rays", ["camera", "random", "axis"])
def test_svo_equals_brute_force_on_tree_leaves(rays):
    """svo over the octree reports the same hit set and depths as brute force
    over the octree's leaf centres (no truncation: hits < n_max)."""
    sc = shell()
    cen, ch = shell_octree()
    leaves = cen[: sc.num_voxels]
    if rays == "camera":
        rs, rd = shell_rays(width=48)
    elif rays == "random":
        rs, rd = random_rays(1500, 1)
    else:
        rs, rd = axis_rays(sc, 1500, 2)
    si, slo, shi = O.svo_intersect(rs[None], rd[None], cen[None], ch[None], sc.voxel_size, 60)
    ai, alo, ahi = O.aabb_intersect(rs[None], rd[None], leaves[None], sc.voxel_size, 60)
    full = (ai[0] != -1).sum(-1) == 60     # truncated: order-dependent, skip
    assert full.mean() < 0.05
    for r in np.nonzero(~full)[0]:
        a = {k: (l, h) for k, l, h in zip(ai[0, r], alo[0, r], ahi[0, r]) if k != -1}
        s = {k: (l, h) for k, l, h in zip(si[0, r], slo[0, r], shi[0, r]) if k != -1}
        assert a == s, r


def test_svo_truncation_follows_dfs_order():
    """With n_max small the svo keeps the first leaves in traversal order
    (child slot 7 first), not in depth or index order."""
    sc = shell()
    cen, ch = shell_octree()
    rs, rd = shell_rays(width=16)
    full = O.svo_intersect(rs[None], rd[None], cen[None], ch[None], sc.voxel_size, 60)[0][0]
    cut = O.svo_intersect(rs[None], rd[None], cen[None], ch[None], sc.voxel_size, 3)[0][0]
    for r in range(rs.shape[0]):
        n = (full[r] != -1).sum()
        assert list(cut[r][: min(n, 3)]) == list(full[r][: min(n, 3)])


# ---------------------------------------------------------------------------
# Octree

def test_octree_two_points_hand_computed():
    c, ch = O.build_octree([1, 1, 1], np.array([[0, 0, 0], [2, 2, 2]]), 0)
    assert c.shape == (3, 3)
    assert list(ch[2]) == [0, -1, -1, -1, -1, -1, -1, 1, 2]
    assert list(c[2]) == [1, 1, 1]
    assert list(ch[0]) == [-1] * 8 + [1] and list(ch[1]) == [-1] * 8 + [1]
    assert list(c[0]) == [0, 0, 0] and list(c[1]) == [2, 2, 2]


def test_octree_depth2_hand_computed():
    c, ch = O.build_octree([3, 0, 0], np.array([[0, 0, 0], [6, 0, 0]]), 2)
    assert c.shape[0] == 7
    assert list(ch[6]) == [5, 4, -1, -1, -1, -1, -1, -1, 8]
    assert list(c[6]) == [3, 0, 0]
    assert list(c[5]) == [1, -2, -2] and list(ch[5]) == [-1] * 6 + [3, -1, 4]
    assert list(c[4]) == [5, -2, -2] and list(ch[4]) == [-1] * 7 + [2, 4]
    assert list(c[3]) == [0, -1, -1] and list(ch[3]) == [-1] * 6 + [0, -1, 2]
    assert list(c[2]) == [6, -1, -1] and list(ch[2]) == [-1] * 6 + [1, -1, 2]


def _pointer_octree(center, points, depth):
    """Independent Python restatement of octree.cpp's insert + BFS finalize."""
    root = {"c": list(center), "d": depth, "i": -1, "ch": [None] * 8}
    for k, p in enumerate(points):
        n = root
        while True:
            bits = [int(p[a] > n["c"][a]) for a in range(3)]
            o = bits[0] + 2 * bits[1] + 4 * bits[2]
            if n["d"] == 0:
                n["ch"][o] = {"c": list(p), "d": -1, "i": k, "ch": [None] * 8}
                break
            if n["ch"][o] is None:
                ln = 1 << (n["d"] - 1)
                n["ch"][o] = {"c": [n["c"][a] + (2 * bits[a] - 1) * ln for a in range(3)],
                              "d": n["d"] - 1, "i": -1, "ch": [None] * 8}
            n = n["ch"][o]
    stack, T = [root], 0
    while stack:
        n = stack.pop()
        T += 1
        stack += [c for c in n["ch"] if c is not None]
    cen = np.zeros((T, 3), np.int32)
    chl = -np.ones((T, 9), np.int32)
    nid = T - 1
    root["i"] = nid
    q = [root]
    while q:
        n = q.pop(0)
        for s, c in enumerate(n["ch"]):
            if c is None:
                continue
            if c["d"] > -1:
                nid -= 1
                c["i"] = nid
            q.append(c)
            chl[n["i"], s] = c["i"]
        chl[n["i"], 8] = 1 << (n["d"] + 1)
        cen[n["i"]] = n["c"]
    return cen, chl


@pytest.mark.parametrize("seed", range(5))
def test_octree_matches_pointer_restatement(seed):
    rng = np.random.default_rng(seed)
    g = int(rng.integers(2, 12))
    ijk = np.argwhere(rng.random((g, g, g)) < 0.3)
    if len(ijk) < 2:
        ijk = np.array([[0, 0, 0], [1, 2, 1]])
    co = (ijk[rng.permutation(len(ijk))] * 2).astype(np.int64)
    center = (co.max(0) + co.min(0)) // 2
    depth = max(int(math.ceil(math.log2((co.max(0) - co.min(0)).max()))) - 1, 0)
    c1, h1 = O.build_octree(center, co, depth)
    c2, h2 = _pointer_octree(center, co, depth)
    assert np.array_equal(c1, c2) and np.array_equal(h1, h2)


def test_octree_shell_known_size_and_invariants():
    sc = shell()
    cen_f, ch = shell_octree()
    N, T = sc.num_voxels, ch.shape[0]
    assert T == 14521
    assert (ch[:N, 8] == 1).all() and (ch[:N, :8] == -1).all()
    assert (ch[N:, 8] > 1).all()
    # every node except the root has exactly one parent; scales halve
    par = {}
    for i in range(N, T):
        for s in range(8):
            c = ch[i, s]
            if c >= 0:
                assert c not in par
                par[c] = i
                assert ch[c, 8] * 2 == ch[i, 8]
    assert len(par) == T - 1 and T - 1 not in par
    # leaves' float centres are the voxel centres
    assert np.abs(cen_f[:N] - sc.points).max() < 1e-6


# ---------------------------------------------------------------------------
# Inverse-CDF sampling

def icdf_one(bins, probs, steps, noise=0.5, compat=True, P=None, extra_rows=None):
    """Run the op on one ray (plus optional following rows) [1,R,P]."""
    P = P or len(bins)
    rows = [(bins, probs, steps)] + (extra_rows or [])
    R = len(rows)
    idx = -np.ones((1, R, P), np.int32)
    lo = np.full((1, R, P), 1e4, np.float32)
    hi = np.full((1, R, P), 1e4, np.float32)
    pr = np.zeros((1, R, P), np.float32)
    st = np.zeros((1, R), np.float32)
    for r, (b, p, s) in enumerate(rows):
        for k, (i, a, z) in enumerate(b):
            idx[0, r, k], lo[0, r, k], hi[0, r, k] = i, a, z
            pr[0, r, k] = p[k]
        st[0, r] = s
    S = int(np.ceil(st).max()) + P
    if np.ndim(noise):
        # per-step noise of the first ray
        nz = np.full((1, R, S), 0.5, np.float32)
        nz[0, 0, :len(noise)] = noise
    else:
        nz = np.full((1, R, S), noise, np.float32)
    oi, od, ox, fl = O.inverse_cdf_sampling_op(idx, lo, hi, nz, pr, st, -1.0, compat)
    n = (oi[0, 0] != -1).sum()
    return oi[0, 0], od[0, 0], ox[0, 0]


def test_icdf_single_bin_closed_form():
    oi, od, ox = icdf_one([(5, 1.0, 2.0)], [1.0], 4.0)
    n = 5
    assert list(oi[:n]) == [5] * 5 and (oi[n:] == -1).all()
    assert list(od[:n]) == [1.0625, 1.25, 1.5, 1.75, 1.9375]
    assert list(ox[:n]) == [0.125, 0.25, 0.25, 0.25, 0.125]


def test_icdf_single_bin_with_noise_closed_form():
    """Step i sits at cdf (i + u_i) / steps (clib/__init__.py:250-256 with
    training-mode noise): z_i = 1 + (i + u_i) / 4 on the bin [1, 2]."""
    u = [0.25, 0.5, 0.75, 0.125]
    oi, od, ox = icdf_one([(5, 1.0, 2.0)], [1.0], 4.0, noise=u)
    z = [1.0] + [1.0 + (i + u[i]) / 4 for i in range(4)] + [2.0]
    assert list(oi[:5]) == [5] * 5 and (oi[5:] == -1).all()
    assert list(od[:5]) == [(z[j] + z[j + 1]) / 2 for j in range(5)]
    assert list(ox[:5]) == [z[j + 1] - z[j] for j in range(5)]


def test_hash_noise_is_uniform_and_keyed():
    """The counter-based noise: the PCG hash is a bijection of uint32 (no
    collisions over 2^20 consecutive inputs); draws are uniform on
    [0.001, 0.999] (mean, quartiles, clamp mass 0.001 per side) and
    decorrelated across steps, rays and seeds; the same key gives the same
    draws."""
    h = O.pcg_hash(np.arange(1 << 20, dtype=np.uint32))
    assert len(np.unique(h)) == 1 << 20
    u = O.hash_noise(3, np.arange(2000), 500)
    assert u.dtype == np.float32 and u.min() >= np.float32(0.001) and u.max() <= np.float32(0.999)
    n = u.size
    assert abs(u.mean() - 0.5) < 5 / np.sqrt(12 * n)
    for q in (0.25, 0.5, 0.75):
        assert abs((u < q).mean() - q) < 5 * np.sqrt(q * (1 - q) / n)
    lo = (u == np.float32(0.001)).mean()
    assert abs(lo - 0.001) < 5 * np.sqrt(0.001 / n)
    c = lambda a, b: abs(np.corrcoef(a.ravel(), b.ravel())[0, 1])
    assert c(u[:, :-1], u[:, 1:]) < 0.01
    assert c(u[:-1], u[1:]) < 0.01
    assert c(u, O.hash_noise(4, np.arange(2000), 500)) < 0.01
    assert np.array_equal(u[7:9], O.hash_noise(3, [7, 8], 500))
    assert O.pass_seed(5, 0) == 5 and O.pass_seed(2**32 - 1, 1) == (2**32 - 1 + 0x9E3779B9) % 2**32


def test_render_view_with_noise_stays_in_bins():
    """With noise the coarse samples still tile each ray's hit segments
    (every sample inside its voxel's segment, the lengths summing to the
    segments' total) and differ from noise 0.5.  (Depths need not increase:
    a zero-length hit sorted after a longer one with the same entry depth
    gets its sample after that one's.)"""
    from _cases import shell, shell_rays
    sc = shell()
    rs, rd = shell_rays(width=16, seed=3)
    a = O.render_view(sc, rs, rd, compat=False, noise_seed=9, hierarchical=False)
    b = O.render_view(sc, rs, rd, compat=False, hierarchical=False)
    assert a["c_idx"].shape[0] == b["c_idx"].shape[0] > 0
    assert not np.array_equal(a["c_depth"], b["c_depth"])
    idx, lo, hi = a["sorted_idx"][a["hits"]], a["sorted_lo"][a["hits"]], a["sorted_hi"][a["hits"]]
    for n in range(idx.shape[0]):
        m = a["c_idx"][n] != -1
        d, x, v = a["c_depth"][n][m], a["c_dist"][n][m], a["c_idx"][n][m]
        seg = {int(i): (l, h) for i, l, h in zip(idx[n], lo[n], hi[n]) if i != -1}
        for dd, xx, vv in zip(d, x, v):
            l, h = seg[int(vv)]
            assert l - 1e-5 <= dd - xx / 2 and dd + xx / 2 <= h + 1e-5
        tot = sum(h - l for l, h in seg.values())
        assert abs(x.astype(np.float64).sum() - tot) < 1e-4 * max(1.0, tot)


def test_icdf_two_bins_with_gap():
    oi, od, ox = icdf_one([(3, 1.0, 2.0), (9, 3.0, 4.0)], [0.5, 0.5], 4.0)
    assert list(oi[:6]) == [3, 3, 3, 9, 9, 9] and (oi[6:] == -1).all()
    assert list(od[:6]) == [1.125, 1.5, 1.875, 3.125, 3.5, 3.875]
    assert list(ox[:6]) == [0.25, 0.5, 0.25, 0.25, 0.5, 0.25]


def test_icdf_full_row_duplicates_next_rays_first_id_in_compat():
    bins = [(3, 1.0, 2.0), (9, 3.0, 4.0)]
    nxt = ([(77, 1.0, 2.0), (78, 3.0, 4.0)], [0.5, 0.5], 4.0)
    oi, od, ox = icdf_one(bins, [0.25, 0.25], 4.0, compat=True, extra_rows=[nxt])
    assert list(oi[:5]) == [3, 3, 9, 9, 77]
    assert list(od[:5]) == [1.25, 1.75, 3.25, 3.75, 3.75]
    oi2, _, _ = icdf_one(bins, [0.25, 0.25], 4.0, compat=False, extra_rows=[nxt])
    assert list(oi2[:4]) == [3, 3, 9, 9] and oi2[4] == -1


def test_icdf_trailing_loop_stops_on_row0_slot_in_compat():
    """The trailing loop's stop test reads ray 0 of the batch row: a longer
    ray behind a one-bin ray 0 loses its last bins in compat mode."""
    ray0 = ([(1, 0.0, 1.0)], [1.0, 0.0, 0.0], 1.0)
    ray1 = ([(4, 1.0, 2.0), (5, 3.0, 4.0), (6, 5.0, 6.0)], [1 / 3, 1 / 3, 1 / 3], 1.0)

    def run(compat):
        rows = [ray0, ray1]
        P = 3
        idx = -np.ones((1, 2, P), np.int32)
        lo = np.full((1, 2, P), 1e4, np.float32)
        hi = np.full((1, 2, P), 1e4, np.float32)
        pr = np.zeros((1, 2, P), np.float32)
        st = np.zeros((1, 2), np.float32)
        for r, (b, p, s) in enumerate(rows):
            for k, (i, a, z) in enumerate(b):
                idx[0, r, k], lo[0, r, k], hi[0, r, k] = i, a, z
            pr[0, r, :] = np.float32(p)
            st[0, r] = s
        nz = np.full((1, 2, 4), 0.5, np.float32)
        return O.inverse_cdf_sampling_op(idx, lo, hi, nz, pr, st, -1.0, compat)

    oi, od, _, _ = run(True)
    assert list(oi[0, 1][:4]) == [4, 5, 5, -1]
    assert list(od[0, 1][:3]) == [1.5, 3.25, 3.75]
    oi, od, _, _ = run(False)
    assert list(oi[0, 1][:4]) == [4, 5, 5, 6]
    assert od[0, 1][3] == 5.5


def test_icdf_invariants_on_camera_rays():
    sc = shell()
    from _cases import hits_for_sampling
    idx, lo, hi, probs, steps, _ = hits_for_sampling(sc, *shell_rays(width=48))
    si, sd, sx, S, fl = O.inverse_cdf_sampling(idx, lo, hi, probs, steps, -1.0, True, compat=False)
    for r in range(idx.shape[0]):
        nb = (idx[r] != -1).sum()
        ns = (si[r] != -1).sum()
        # one tail per bin, plus one per CDF step (k + 0.5) / steps that stays
        # inside the total probability 1, i.e. round(steps) of them
        frac = float(steps[r]) - math.floor(float(steps[r]))
        if abs(frac - 0.5) > 1e-4:
            assert ns == nb + int(math.floor(float(steps[r]) + 0.5))
        seg = float((hi[r, :nb] - lo[r, :nb]).astype(np.float64).sum())
        assert abs(float(sx[r, :ns].astype(np.float64).sum()) - seg) < 1e-5   # length conserved
        assert (np.diff(sd[r, :ns]) > 0).all()             # depth-ordered
        # every sample lies inside its voxel's interval
        for k in range(ns):
            j = list(idx[r, :nb]).index(si[r, k])
            assert lo[r, j] - 1e-6 <= sd[r, k] <= hi[r, j] + 1e-6


# ---------------------------------------------------------------------------
# Uniform sampling

def uniform_one(bins, step, S=32, noise=0.5):
    P = len(bins) + 1
    idx = -np.ones((1, 1, P), np.int32)
    lo = np.full((1, 1, P), 1e4, np.float32)
    hi = np.full((1, 1, P), 1e4, np.float32)
    for k, (i, a, z) in enumerate(bins):
        idx[0, 0, k], lo[0, 0, k], hi[0, 0, k] = i, a, z
    nz = np.full((1, 1, S), noise, np.float32)
    oi, od, ox, fl = O.uniform_ray_sampling_op(idx, lo, hi, nz, step, S)
    return oi[0, 0], od[0, 0], ox[0, 0]


def test_uniform_single_bin_closed_form():
    oi, od, ox = uniform_one([(5, 1.0, 2.0)], 0.25)
    assert list(oi[:5]) == [5] * 5 and (oi[5:] == -1).all()
    assert list(od[:5]) == [1.0625, 1.25, 1.5, 1.75, 1.9375]
    assert list(ox[:5]) == [0.125, 0.25, 0.25, 0.25, 0.125]


def test_uniform_skips_the_gap():
    oi, od, ox = uniform_one([(3, 1.0, 2.0), (9, 3.0, 4.0)], 0.5)
    assert list(oi[:6]) == [3, 3, 3, 9, 9, 9] and (oi[6:] == -1).all()
    assert list(od[:6]) == [1.125, 1.5, 1.875, 3.125, 3.5, 3.875]
    assert list(ox[:6]) == [0.25, 0.5, 0.25, 0.25, 0.5, 0.25]


# ---------------------------------------------------------------------------
# Glue, gather, compositing, prune, split

def test_probs_and_steps():
    idx = np.array([[1, 2, -1]], np.int32)
    lo = np.array([[1.0, 3.0, 1e4]], np.float32)
    hi = np.array([[2.0, 3.5, 1e4]], np.float32)
    probs, steps = O.probs_steps(idx, lo, hi, 0.25)
    assert list(probs[0]) == [np.float32(1 / 1.5), np.float32(0.5 / 1.5), 0.0]
    assert steps[0] == 6.0


def test_sort_hits_pads_last_and_stable():
    idx = np.array([[4, -1, 2, 7]], np.int32)
    lo = np.array([[2.0, 0.0, 1.0, 2.0]], np.float32)
    hi = np.array([[3.0, 0.0, 2.0, 3.5]], np.float32)
    i, a, b, h = O.sort_hits(idx, lo, hi)
    assert list(i[0]) == [2, 4, 7, -1] and list(a[0]) == [1, 2, 2, 1e4] and h[0]


def test_trilinear_reproduces_affine_fields():
    sc = shell()
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 5))
    b0 = rng.standard_normal(5)
    # corner key k sits at key * half_voxel + offset (same frame as centres)
    coords, res = O.discretize_points(sc.points, np.float32(sc.voxel_size) / 2)
    key_pos = sc.keys.astype(np.float64) * (sc.voxel_size / 2) + res.astype(np.float64)
    values = key_pos @ A + b0
    cen = coords.astype(np.float64) * (sc.voxel_size / 2) + res.astype(np.float64)
    vi = rng.integers(0, sc.num_voxels, 500)
    x = cen[vi] + rng.uniform(-0.5, 0.5, (500, 3)) * sc.voxel_size
    emb = O.trilinear_gather(vi, x, cen, sc.feats, values, sc.voxel_size)
    assert np.abs(emb - (x @ A + b0)).max() < 1e-10
    W = O.trilinear_weights(vi, x, cen, sc.voxel_size)
    assert np.abs(W.sum(1) - 1).max() < 1e-12


def test_trilinear_corner_order_matches_feats():
    """At a voxel's corner c (4x + 2y + z order) the gather returns exactly
    that corner's embedding."""
    sc = shell()
    coords, res = O.discretize_points(sc.points, np.float32(sc.voxel_size) / 2)
    cen = coords.astype(np.float64) * (sc.voxel_size / 2)
    v = 123
    for c in range(8):
        off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1]) - 0.5
        x = cen[v] + off * sc.voxel_size
        emb = O.trilinear_gather([v], x[None], cen, sc.feats, sc.values, sc.voxel_size)
        assert np.abs(emb[0] - sc.values[sc.feats[v, c]]).max() < 1e-12


def test_composite_constant_sigma_closed_form():
    R, S = 3, 10
    sigma = np.full((R, S), 2.0)
    dists = np.full((R, S), 0.01)
    idx = np.zeros((R, S), np.int32)
    idx[1, 5:] = -1
    depth = np.tile(np.arange(S, dtype=np.float64), (R, 1))
    rgb = np.ones((R, S, 3)) * 0.25
    col, d, missed, probs = O.composite(sigma, rgb, depth, dists, idx)
    fe = 2.0 * 0.01 * 7.0
    assert abs(missed[0] - math.exp(-fe * S)) < 1e-12
    assert abs(missed[1] - math.exp(-fe * 5)) < 1e-12
    assert abs(probs[0, 0] - (1 - math.exp(-fe))) < 1e-12
    assert abs(probs[0, 3] - (1 - math.exp(-fe)) * math.exp(-3 * fe)) < 1e-12
    assert np.abs(probs.sum(1) + missed - 1).max() < 1e-12
    assert np.abs(col[0] - 0.25 * (1 - missed[0])).max() < 1e-12


def _composite_objective(sigma, rgb, depth, dists, idx, g_rgb, g_depth, g_missed):
    col, d, missed, _ = O.composite(sigma, rgb, depth, dists, idx)
    return (col * g_rgb).sum() + (d * g_depth).sum() + (missed * g_missed).sum()


def test_composite_backward_matches_central_differences():
    """Every sigma and colour gradient of the FP64 restatement against central
    differences of composite() (pads, negative sigmas and zero-length
    intervals included)."""
    rng = np.random.default_rng(11)
    R, S = 5, 9
    sigma = rng.normal(1.0, 2.0, (R, S))
    dists = rng.uniform(0.0, 0.2, (R, S))
    dists[0, 2] = 0.0
    idx = np.zeros((R, S), np.int32)
    idx[1, 6:] = -1
    idx[3, 1:] = -1
    depth = np.cumsum(rng.uniform(0.1, 0.3, (R, S)), 1)
    rgb = rng.uniform(0, 1, (R, S, 3))
    g_rgb, g_depth, g_missed = rng.normal(size=(R, 3)), rng.normal(size=R), rng.normal(size=R)
    gs, gc = O.composite_backward(sigma, rgb, depth, dists, idx, g_rgb, g_depth, g_missed)
    f = lambda sg, c: _composite_objective(sg, c, depth, dists, idx, g_rgb, g_depth, g_missed)
    eps = 1e-6
    for r in range(R):
        for k in range(S):
            if abs(sigma[r, k]) < 10 * eps:
                continue
            a, b = sigma.copy(), sigma.copy()
            a[r, k] += eps
            b[r, k] -= eps
            fd = (f(a, rgb) - f(b, rgb)) / (2 * eps)
            assert abs(fd - gs[r, k]) < 1e-7 * max(1.0, abs(fd)), (r, k, fd, gs[r, k])
            for ch in range(3):
                a, b = rgb.copy(), rgb.copy()
                a[r, k, ch] += eps
                b[r, k, ch] -= eps
                fd = (f(sigma, a) - f(sigma, b)) / (2 * eps)
                assert abs(fd - gc[r, k, ch]) < 1e-7 * max(1.0, abs(fd))
    assert (gs[idx == -1] == 0).all() and (gs[(sigma <= 0)] == 0).all()


def test_composite_backward_single_sample_closed_form():
    """One sample: rgb = (1 - exp(-7 s d)) c, so d rgb.g / ds = 7 d exp(-7 s d) <c, g>."""
    s, d = 1.7, 0.05
    c = np.array([[[0.2, 0.5, 0.9]]])
    g = np.array([[0.3, -1.0, 2.0]])
    gs, gc = O.composite_backward(np.array([[s]]), c, np.array([[1.0]]), np.array([[d]]),
                                  np.zeros((1, 1), np.int32), g, np.zeros(1), np.zeros(1))
    assert abs(gs[0, 0] - 7 * d * math.exp(-7 * s * d) * float(c[0, 0] @ g[0])) < 1e-14
    assert np.abs(gc[0, 0] - (1 - math.exp(-7 * s * d)) * g[0]).max() < 1e-15


def test_trilinear_backward_is_the_adjoint_of_the_gather():
    """<gather(V), G> == <V, gather_backward(G)> for random V and G (the
    backward is exactly the transpose of the forward's linear map), and a
    sample at a voxel centre spreads 1/8 of its gradient to each corner."""
    sc = shell()
    rng = np.random.default_rng(5)
    cen = O.shift_centers(sc.points, sc.voxel_size).astype(np.float64)
    M, D = 700, 6
    vi = rng.integers(0, sc.num_voxels, M)
    vi[::17] = -1
    x = cen[np.maximum(vi, 0)] + rng.uniform(-0.5, 0.5, (M, 3)) * sc.voxel_size
    V = rng.standard_normal((sc.num_keys, D))
    G = rng.standard_normal((M, D))
    ok = vi >= 0
    emb = O.trilinear_gather(vi[ok], x[ok], cen, sc.feats, V, sc.voxel_size)
    gv = O.trilinear_backward(vi, x, cen, sc.feats, sc.num_keys, G, sc.voxel_size)
    lhs = (emb * G[ok]).sum()
    rhs = (V * gv).sum()
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))
    v = 42
    g1 = O.trilinear_backward([v], cen[v][None], cen, sc.feats, sc.num_keys, np.ones((1, 2)),
                              sc.voxel_size)
    assert np.allclose(g1[sc.feats[v]], 0.125, atol=1e-15)
    assert abs(g1.sum() - 2.0) < 1e-14


def test_fixed_sample_loss_restates_render_view():
    """The FP64 training objective with samples held fixed equals the squared
    error of render_view's image (plus the constant of the rays that miss),
    in full and over a subset of rays."""
    sc = shell()
    rs, rd = shell_rays(width=24, seed=3)
    ref = O.render_view(sc, rs, rd, hierarchical=False)
    tgt = np.random.default_rng(1).uniform(0, 1, (rs.shape[0], 3))
    obj = O.FixedSampleLoss(ref["c_idx"], ref["c_depth"], ref["c_dist"], ref["hits"], rs, rd,
                            O.shift_centers(sc.points, sc.voxel_size), sc.feats, sc.voxel_size,
                            tgt)
    miss = ~ref["hits"]
    err = ((ref["rgb"] - tgt) ** 2).sum(1)
    assert abs(obj.loss(sc.values, sc.mlp) + err[miss].sum() - err.sum()) < 1e-9 * err.sum()
    rows = [0, 3, 5]
    frame = np.nonzero(ref["hits"])[0][rows]
    assert abs(obj.loss(sc.values, sc.mlp, rows) - err[frame].sum()) < 1e-12 * err.sum()
    k = int(obj.keys[obj.rows == 3][0, 0])
    assert 3 in obj.rows_touching_key(k)


def test_mlp_central_difference_equals_two_loss_evaluations():
    """The rank-1 shortcut for MLP-entry central differences gives the same
    number as perturbing the MLP and evaluating the loss twice."""
    sc = shell()
    rs, rd = shell_rays(width=16, seed=4)
    ref = O.render_view(sc, rs, rd, hierarchical=False)
    tgt = np.random.default_rng(2).uniform(0, 1, (rs.shape[0], 3))
    obj = O.FixedSampleLoss(ref["c_idx"], ref["c_depth"], ref["c_dist"], ref["hits"], rs, rd,
                            O.shift_centers(sc.points, sc.voxel_size), sc.feats, sc.voxel_size,
                            tgt)
    emb = obj.embed(sc.values)
    eps = 1e-4
    rng = np.random.default_rng(0)
    for layer in range(len(sc.mlp.weights)):
        W = sc.mlp.weights[layer]
        for kind, pos in (("w", (int(rng.integers(W.shape[0])), int(rng.integers(W.shape[1])))),
                          ("b", (int(rng.integers(W.shape[0])),))):
            def at(delta):
                ws = [w.astype(np.float64) for w in sc.mlp.weights]
                bs = [b.astype(np.float64) for b in sc.mlp.biases]
                (ws if kind == "w" else bs)[layer][pos] += delta
                return obj.loss_from_embedding(emb, O.MLPWeights(ws, bs))
            naive = (at(eps) - at(-eps)) / (2 * eps)
            fast = obj.mlp_central_difference(emb, sc.mlp, layer, kind, pos, eps)
            assert abs(fast - naive) <= 1e-9 * max(1e-3, abs(naive)), (layer, kind, fast, naive)


def test_prune_lattice_includes_corners():
    pts = np.zeros((1, 3), np.float32)
    lat = O.prune_lattice(pts, 0.5, bits=16)
    assert lat.shape == (1, 4096, 3)
    assert list(lat[0, 0]) == [-0.25, -0.25, -0.25] and list(lat[0, -1]) == [0.25, 0.25, 0.25]


def test_split_reproduces_parent_corner_values():
    sc = shell()
    sub = np.arange(0, sc.num_voxels, 37)
    pts = sc.points[sub]
    keys, feats = O.voxel_keys(pts, sc.voxel_size)
    rng = np.random.default_rng(0)
    values = rng.standard_normal((len(keys), 4))
    np_, nf, nv, nk = O.split_voxels(pts, feats, values, sc.voxel_size)
    assert np_.shape == (8 * len(pts), 3) and nf.shape == (8 * len(pts), 8)
    assert (np.diff(nk[:, 0]) >= 0).all()
    # a child's outer corner coincides with its parent's corner: same value
    for v in range(len(pts)):
        for c in range(8):
            child = 8 * v + c
            assert np.abs(nv[nf[child, c]] - values[feats[v, c]]).max() < 1e-12
    # children at parent +- vs/4
    d = np_.reshape(-1, 8, 3) - pts[:, None, :]
    assert np.allclose(np.abs(d), sc.voxel_size / 4)


def test_prune_max_sigma_constant_field():
    """A field whose sigma is a constant s (all weights zero, sigma bias s)
    gives max relu(sigma) = max(s, 0) for every voxel; and with a sigma linear
    in one embedding feature that is affine in position, the max sits at the
    lattice corner (the lattice includes the corners)."""
    sc = shell()
    cen = O.shift_centers(sc.points, sc.voxel_size)
    D = sc.values.shape[1]
    ws = [np.zeros((4, D))]
    for s in (-0.3, 1.25):
        bs = [np.array([s, 0.0, 0.0, 0.0])]
        m = O.prune_max_sigma(cen, sc.feats, sc.values, sc.voxel_size, O.MLPWeights(ws, bs),
                              [0, 7, 100], bits=4)
        assert np.allclose(m, max(s, 0.0), atol=0, rtol=0)
    # values = x coordinate of each corner key (affine field): sigma = emb[0]
    coords, res = O.discretize_points(sc.points, np.float32(sc.voxel_size) / 2)
    key_pos = sc.keys.astype(np.float64) * (sc.voxel_size / 2) + res.astype(np.float64)
    vals = np.zeros((sc.num_keys, D))
    vals[:, 0] = key_pos[:, 0] + 5.0
    w1 = np.zeros((4, D))
    w1[0, 0] = 1.0
    m = O.prune_max_sigma(cen, sc.feats, vals, sc.voxel_size,
                          O.MLPWeights([w1], [np.zeros(4)]), [3, 50], bits=8)
    # max x over the lattice = centre_x + vs/2 in the key frame
    ctr = coords[[3, 50], 0].astype(np.float64) * (sc.voxel_size / 2) + res[0, 0]
    assert np.abs(m - (ctr + sc.voxel_size / 2 + 5.0)).max() < 1e-5


def test_fine_normaliser_order():
    """The fine-pass normaliser: exact on values whose sums are exact, equal
    to a hand-evaluated strided-then-butterfly order on a small case, and
    within rounding of the float64 sum."""
    x = np.full((1, 200), 0.25, np.float32)
    assert O.fine_normaliser(x)[0] == np.float32(50.0)
    rng = np.random.default_rng(0)
    y = rng.uniform(0, 1, (3, 70)).astype(np.float32)
    part = [np.float32(0)] * 64
    for k in range(70):
        part[k % 64] = np.float32(part[k % 64] + y[1, k])
    for d in (1, 2, 4, 8, 16, 32):
        part = [np.float32(part[j] + part[j ^ d]) for j in range(64)]
    assert O.fine_normaliser(y)[1] == part[0]
    assert np.abs(O.fine_normaliser(y) - y.astype(np.float64).sum(1)).max() < 1e-5


# ---- camera rays (renderer.py:91-112, reader.py:92-119, geometry.py:76-122)

def test_camera_uv_enumerates_pixels():
    """get_uv: u = column * W/w, v = row * H/h, as a float32 [2, h, w]."""
    uv = O.camera_uv(9, 4, 3, 2)      # rows scaled by 3, columns by 2
    assert uv.dtype == np.float32 and uv.shape == (2, 3, 2)
    for i in range(3):
        for j in range(2):
            assert uv[0, i, j] == np.float32(j * 2.0) and uv[1, i, j] == np.float32(i * 3.0)


def test_camera_rays_closed_forms():
    """Identity pose, f = 1, principal point 0: pixel (i, j) looks along
    (j, i, 1) / |.|; the principal-point pixel of any pose looks along the
    third column of R; directions are unit; starts are the translation."""
    rs, rd = O.camera_rays(3, 4, 3, 4, 1.0, 1.0, 0.0, 0.0, np.eye(4))
    i, j = np.divmod(np.arange(12), 4)
    ref = np.stack([j, i, np.ones(12)], 1).astype(np.float64)
    ref /= np.linalg.norm(ref, axis=1, keepdims=True) + 1e-8     # normalize(): |d| + 1e-8
    assert np.abs(rd - ref).max() < 1e-12 and not rs.any()
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    RT = np.eye(4)
    RT[:3, :3] = q
    RT[:3, 3] = [0.5, -2.0, 1.5]
    rs, rd = O.camera_rays(8, 10, 8, 10, 7.0, 7.0, 5.0, 4.0, RT)
    assert np.abs(rd[4 * 10 + 5] - q[:, 2] / (1 + 1e-8)).max() < 1e-12
    assert np.abs(np.linalg.norm(rd, axis=1) - 1).max() < 1e-7
    assert np.array_equal(rs, np.broadcast_to(RT[:3, 3], rs.shape))


def test_camera_rays_match_the_scene_cameras():
    """Against the scenes' own look-at camera (an independent float64
    restatement: R [(u - cx)/f, (v - cy)/f, 1], normalised)."""
    from nsvf_amd.rays import ViewCamera
    cam_pos = S.sphere_camera(2.5, 11)
    rs0, rd0 = S.camera_rays(40, 30, cam_pos)
    vc = ViewCamera.looking_at(cam_pos, 30, 40)
    fx, fy, cx, cy = vc.intrinsics[0, 0], vc.intrinsics[1, 1], vc.intrinsics[0, 2], vc.intrinsics[1, 2]
    rs, rd = O.camera_rays(30, 40, 30, 40, fx, fy, cx, cy, vc.inv_RT)
    assert np.abs(rd - rd0).max() < 1e-6 and np.abs(rs - rs0).max() < 1e-6


# ---- train-set voxel statistics (encoder.py:594-603)

def test_track_voxel_probs_closed_form():
    """A voxel's samples on one ray add up (also when split into two runs),
    rays take the max, -1 samples are dropped, and the running values only
    grow."""
    idx = np.array([[3, 3, 5, 3, -1], [5, 1, -1, -1, -1]])
    p = np.array([[0.1, 0.2, 0.4, 0.3, 0.9], [0.5, 0.05, 0.7, 0.0, 0.0]], np.float32)
    out = O.track_voxel_probs(idx, p, 6)
    s3 = np.float32(np.float32(np.float32(0.1) + np.float32(0.2)) + np.float32(0.3))
    assert out[3] == s3 and out[5] == np.float32(0.5) and out[1] == np.float32(0.05)
    assert out[0] == 0 and out[2] == 0 and out[4] == 0
    prev = np.array([0, 0.01, 0, 0.7, 0, 0.2], np.float32)
    out2 = O.track_voxel_probs(idx, p, 6, prev)
    assert out2[3] == np.float32(0.7) and out2[5] == np.float32(0.5) and out2[1] == np.float32(0.05)


def test_track_voxel_probs_matches_scatter_add():
    """Against an independent form: numpy's unbuffered in-order add.at per
    ray into N + 1 columns (the reference's dropped -1 column), then max."""
    rng = np.random.default_rng(3)
    N, R, S = 40, 30, 25
    idx = rng.integers(-1, N, (R, S))
    p = rng.random((R, S)).astype(np.float32)
    ref = np.zeros(N, np.float32)
    for r in range(R):
        col = np.zeros(N + 1, np.float32)
        np.add.at(col, np.where(idx[r] < 0, N, idx[r]), p[r])
        ref = np.maximum(ref, col[:N])
    assert np.array_equal(O.track_voxel_probs(idx, p, N), ref)


# ---- the inverse-CDF wrapper's grouping (clib/__init__.py:234-270): which
# ray is "the next ray" (sample_gpu.cu:188) and "ray 0" (:194) in compat mode

def _wrapper_rows(N, P, counts, probs, steps):
    """N rays; ray r's slot k holds voxel id 1000 r + k for k < counts[r]
    over unit bins [k, k + 1], probabilities probs[r] per slot."""
    idx = np.full((N, P), -1, np.int32)
    lo = np.full((N, P), 1e4, np.float32)
    hi = np.full((N, P), 1e4, np.float32)
    pr = np.zeros((N, P), np.float32)
    for r in range(N):
        c = counts[r]
        idx[r, :c] = 1000 * r + np.arange(c)
        lo[r, :c] = np.arange(c)
        hi[r, :c] = np.arange(c) + 1
        pr[r, :c] = probs[r]
    return idx, lo, hi, pr, np.full(N, np.float32(steps), np.float32)


def _next_ray_by_hand(N):
    """The next ray each ray reads, derived from the wrapper: rows padded to
    H = 200 ceil(N / 200) with copies of row 0 (a pad row's ids are row
    0's), regrouped to [200, H / 200, P] (batch row b holds rays
    (H / 200) b ...), cut into chunks of 800 columns, each a contiguous
    [200, cols, P] buffer: the next ray of (b, c) is (b, c + 1) inside the
    chunk, the first column of the chunk in batch row b + 1 after its last
    column, and past the end of the buffer (-1) for the last ray of the
    chunk.  Returns the ray whose ids the next slot holds (-1: none)."""
    G = 200
    H = -(-N // G) * G
    K = H // G
    out = []
    for r in range(N):
        b, c = divmod(r, K)
        c0 = (c // 800) * 800
        width = min(800, K - c0)
        if c + 1 < c0 + width:
            nb, nc = b, c + 1
        elif b + 1 < G:
            nb, nc = b + 1, c0
        else:
            out.append(-1)
            continue
        nr = nb * K + nc
        out.append(nr if nr < N else 0)     # pad rows copy row 0
    return np.array(out)


def _ray0_by_hand(N):
    """Ray 0 of each ray's batch row inside its chunk: (b, first column of
    the chunk)."""
    K = -(-N // 200)
    r = np.arange(N)
    b, c = r // K, r % K
    return b * K + (c // 800) * 800


@pytest.mark.parametrize("N", [401, 200 * 801])
def test_icdf_wrapper_next_ray_is_hand_derived(N):
    """Full rows whose probabilities sum to 0.8 with 4 steps: the last CDF
    step (0.875) runs past the last bin, and in compat mode the trailing
    loop then adds one sample carrying the next slot's voxel id -- the next
    ray's first id, or none (-1) past the buffer."""
    P = 4
    idx, lo, hi, pr, st = _wrapper_rows(N, P, [P] * N, [0.2] * N, 4.0)
    si, _, _, _, _ = O.inverse_cdf_sampling(idx, lo, hi, pr, st, -1.0, True, compat=True)
    nxt = _next_ray_by_hand(N)
    # steps at CDF 0.125 (bin 0), 0.375 (bin 1), 0.625 (bin 3), 0.875 (past
    # bin 3) with the tails of bins 0..3 between them: 7 samples of the ray's
    # own ids, then the extra one
    last = si[:, 7]
    assert (si[:, :7] // 1000 == np.arange(N)[:, None]).all()
    assert np.array_equal(si[0, :7] % 1000, [0, 0, 1, 1, 2, 3, 3])
    want = np.where(nxt >= 0, 1000 * np.maximum(nxt, 0), -1)
    assert np.array_equal(last, want)
    if N == 401:        # the last real ray reads the first pad row, a copy of row 0
        assert last[400] == 0 and last[399] == 400 * 1000
    else:               # a chunk's last column: the next batch row, not the next ray
        assert last[799] == 801 * 1000 and last[800] == 1601 * 1000
    si2, _, _, _, _ = O.inverse_cdf_sampling(idx, lo, hi, pr, st, -1.0, True, compat=False)
    assert (si2[:, 7:] == -1).all()


@pytest.mark.parametrize("N", [401, 200 * 801])
def test_icdf_wrapper_ray0_is_hand_derived(N):
    """Rays of 6 bins whose first bin holds all the probability: the 4 steps
    land in bin 0 and the trailing loop emits bin tails 0, 1, ... until the
    slot after the current one is empty in ray 0 of the batch row (compat
    mode), so a ray gets 4 + min(6, count of its ray 0) samples."""
    P = 8
    counts = np.full(N, 6)
    r0 = _ray0_by_hand(N)
    heads = np.unique(r0)
    counts[heads] = 1 + (np.arange(len(heads)) % 5)       # ray 0s: 1 .. 5 bins
    probs = [[1.0] + [0.0] * 7 for _ in range(N)]
    idx, lo, hi, pr, st = _wrapper_rows(N, P, counts, [np.array(p[:c] + [0] * (P - c))[: c]
                                                        for p, c in zip(probs, counts)], 4.0)
    pr[:, 0] = 1.0
    si, _, _, _, _ = O.inverse_cdf_sampling(idx, lo, hi, pr, st, -1.0, True, compat=True)
    got = (si != -1).sum(1)
    want = 4 + np.minimum(counts, counts[r0])
    assert np.array_equal(got, want)
    si2, _, _, _, _ = O.inverse_cdf_sampling(idx, lo, hi, pr, st, -1.0, True, compat=False)
    assert np.array_equal((si2 != -1).sum(1), 4 + counts)


# ---- hierarchical bins (nerf.py:64-79): [z - d/2, z + d/2], probs
# (p + 1e-5) / sum over every column, pads included

def _two_equal_bins():
    si = np.array([[7, 8, -1, -1]], np.int32)
    sd = np.array([[1.0, 1.5, 1e4, 1e4]], np.float32)
    sx = np.array([[0.5, 0.5, 0.0, 0.0]], np.float32)
    pr = np.array([[0.25, 0.25, 0.0, 0.0]], np.float32)
    return si, sd, sx, pr


def test_fine_bins_hand_computed():
    si, sd, sx, pr = _two_equal_bins()
    fi, flo, fhi, fp, fst = O.prepare_fine_bins(si, sd, sx, pr, 64.0)
    assert np.array_equal(fi, si)
    assert list(flo[0]) == [0.75, 1.25, 1e4, 1e4] and list(fhi[0]) == [1.25, 1.75, 1e4, 1e4]
    den = 0.5 + 4e-5                                 # two samples + 1e-5 on all four columns
    want = np.array([0.25001, 0.25001, 1e-5, 1e-5]) / den
    assert np.abs(fp[0] - want).max() < 2e-7 * want.max()
    assert fst[0] == 64.0
    # the alternatives a slip would give differ by ~2e-5 in p0
    for alt in (0.25001 / (0.5 + 4e-8 * 0 + 2e-5),    # pads left out of the sum
                0.25000001 / 0.50000004):             # the commented-out 1e-8
        assert abs(fp[0, 0] - alt) > 1e-5


def test_fine_samples_hand_computed():
    """64 steps over the two bins: steps k < 32 land in bin 0 at
    z = 0.75 + 0.5 (k + 1/2) / 64 / p0, then bin 0's tail, steps 32..63 in
    bin 1 at z = 1.25 + 0.5 ((k + 1/2) / 64 - p0) / p1, then bin 1's tail:
    66 samples, each (midpoint of z and the previous z, their distance)."""
    si, sd, sx, pr = _two_equal_bins()
    fi, flo, fhi, fp, fst = O.prepare_fine_bins(si, sd, sx, pr, 64.0)
    for compat in (True, False):
        oi, od, ox, _, _ = O.inverse_cdf_sampling(fi, flo, fhi, fp, fst, -1.0, True, compat=compat)
        valid = oi[0] != -1
        assert valid.sum() == 66
        p0 = 0.25001 / 0.50004
        z = [0.75 + 0.5 * (k + 0.5) / 64 / p0 for k in range(32)] + [1.25]
        z += [1.25 + 0.5 * ((k + 0.5) / 64 - p0) / p0 for k in range(32, 64)] + [1.75]
        prev = [0.75] + z[:-1]
        want_d = (np.array(z) + np.array(prev)) / 2
        want_x = np.array(z) - np.array(prev)
        assert list(oi[0][valid]) == [7] * 33 + [8] * 33
        assert np.abs(od[0][valid] - want_d).max() < 1e-6
        assert np.abs(ox[0][valid] - want_x).max() < 1e-6
        # with p0 = 0.5 (1e-8, or pads left out of the sum) step 31 moves by ~2e-5
        z31_alt = 0.75 + 0.5 * 31.5 / 64 / 0.5
        assert abs(z[31] - z31_alt) > 1e-5


def test_fine_bins_reject_swapped_halves():
    si, sd, sx, pr = _two_equal_bins()
    _, flo, fhi, _, _ = O.prepare_fine_bins(si, sd, sx, pr, 64.0)
    assert (flo[0, :2] < fhi[0, :2]).all() and (flo[0, :2] == sd[0, :2] - sx[0, :2] / 2).all()


def test_fill_background_hand_computed():
    """nsvf.py:89-101: hit rays keep their colour plus missed x background
    and their depth plus missed x 5; rays that hit nothing are background
    (rgb 1, depth 5, missed 1)."""
    hits = np.array([False, True, False, True])
    rgb = np.array([[0.1, 0.2, 0.3], [0.0, 0.5, 1.0]])
    depth = np.array([2.0, 3.0])
    missed = np.array([0.25, 0.0])
    frgb, fd, fm = O.fill_background(hits, rgb, depth, missed, bg=(1.0, 0.5, 0.0))
    assert np.allclose(frgb, [[1.0, 0.5, 0.0], [0.35, 0.325, 0.3], [1.0, 0.5, 0.0], [0.0, 0.5, 1.0]],
                       rtol=0, atol=1e-15)
    assert np.allclose(fd, [5.0, 2.0 + 0.25 * 5.0, 5.0, 3.0], rtol=0, atol=1e-15)
    assert np.array_equal(fm, [1.0, 0.25, 1.0, 0.0])


# ---- FP64 slab test and grazing report

def test_aabb_hits_f64_matches_numpy_restatement():
    sc = shell()
    rs, rd = shell_rays(width=16)
    pts = O.shift_centers(sc.points, sc.voxel_size)
    i_np, lo_np, hi_np = O.aabb_intersect_f64(rs, rd, pts, sc.voxel_size, 60)
    i_c, lo_c, hi_c, most = O.aabb_hits_f64(rs, rd, pts, sc.voxel_size, 60)
    assert most <= 60
    assert np.array_equal(i_np, i_c)
    assert np.array_equal(lo_np, lo_c) and np.array_equal(hi_np, hi_c)


def test_grazing_report_counts_a_constructed_graze():
    """Rows that equal the FP64 test are no difference; an FP32 hit the FP64
    test misses counts as a graze when the FP32 crossing is shorter than
    eps, and not when it is long; an FP64 hit missing from the FP32 row is
    judged by its FP64 crossing."""
    pts = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0], [0.0, 20.0, 0.0]], np.float32)
    rs = np.array([[-2.0, 0.0, 0.0], [-2.0, 5.0, 0.0]], np.float32)
    rd = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]], np.float32)
    idx, lo, hi = O.aabb_intersect(rs[None], rd[None], pts[None], 1.0, 4)
    idx, lo, hi = idx[0], lo[0], hi[0]
    assert list(idx[0, :3]) == [0, 1, -1] and (idx[1] == -1).all()
    rep = O.grazing_report(idx, lo, hi, rs, rd, pts, 1.0, 4)
    assert rep["rays"] == 2 and rep["rays_differing_from_fp64"] == 0
    assert rep["max_abs_hit_depth_vs_fp64"] == 0.0

    def case(k, a, b):
        i2, l2, h2 = idx.copy(), lo.copy(), hi.copy()
        i2[0, 2], l2[0, 2], h2[0, 2] = k, np.float32(a), np.float32(b)
        return O.grazing_report(i2, l2, h2, rs, rd, pts, 1.0, 4)

    rep = case(2, 4.5, 4.5)             # FP32 touches box 2 over length 0
    assert rep["rays_differing_from_fp64"] == 1 and rep["differences_that_graze"] == 1
    assert rep["rays_with_hits_shorter_than_eps"] == 1
    rep = case(2, 1.0, 2.0)             # ... over length 1: not a graze
    assert rep["rays_differing_from_fp64"] == 1 and rep["differences_that_graze"] == 0
    i2 = idx.copy()
    i2[0, 1] = -1                       # FP32 drops box 1, which FP64 crosses over 1
    rep = O.grazing_report(i2, lo, hi, rs, rd, pts, 1.0, 4)
    assert rep["rays_differing_from_fp64"] == 1 and rep["differences_that_graze"] == 0


def test_render_view_subset_equals_full_view():
    """Field and compositing on a subset of rays (geometry and sampling over
    all of them) give the full view's outputs at those rays, in both
    sampler modes."""
    sc = shell()
    rs, rd = shell_rays(width=32)
    sel = np.sort(np.random.default_rng(3).choice(rs.shape[0], 300, replace=False))
    for compat in (True, False):
        full = O.render_view(sc, rs, rd, compat=compat)
        sub = O.render_view(sc, rs, rd, compat=compat, subset=sel)
        for k in ("rgb", "depth", "missed"):
            assert np.abs(sub[k] - full[k][sel]).max() < 1e-12
        assert np.array_equal(sub["c_idx"], full["c_idx"]) and np.array_equal(sub["f_idx"][sub["eval_rows"]],
                                                                            full["f_idx"][sub["eval_rows"]])


# ---------------------------------------------------------------------------
# Early ray termination (renderer.py:136-174): chunked marching

def _march_case(counts, hot, S=None):
    """idx [N,S] with counts[n] leading samples, sigma 1 at (row, col) in
    `hot` (else 0), dists 1/7: free energy = sigma."""
    S = S or max(counts)
    idx = -np.ones((len(counts), S), np.int32)
    for n, c in enumerate(counts):
        idx[n, :c] = 1
    sig = np.zeros(idx.shape)
    for n, c in hot:
        sig[n, c] = 1.0
    return idx, sig, np.full(idx.shape, np.float32(1 / 7))


def test_march_stops_at_chunk_ends():
    """Two rows of 6 samples, row 0 past the threshold (-log 0.5) at column
    2: it is evaluated to the end of the chunk holding column 2 -- chunks of
    2 columns at chunk_size 4, of 3 at 6, of 1 at 2."""
    idx, sig, dx = _march_case([6, 6], [(0, 2)])
    for chunk, last in ((4, 3), (6, 2), (2, 2)):
        e, stop, _ = O.march_chunks(idx, sig, dx, 0.5, chunk)
        assert list(e[0]) == [c <= last for c in range(6)], chunk
        assert e[1].all() and list(stop) == [True, False]


def test_march_stopped_rows_leave_the_chunk_counts():
    """Three rows of 4 samples at chunk_size 4: the first chunk is column 0
    (3 hits; two columns would be 6).  Row 0 stops there, so the next chunk
    holds columns 1-2 (2 + 2 hits): row 1, past the threshold at column 1,
    is still evaluated at column 2.  Were stopped rows still counted, the
    chunk would be column 1 alone and row 1 would stop before column 2."""
    idx, sig, dx = _march_case([4, 4, 4], [(0, 0), (1, 1)])
    e, stop, _ = O.march_chunks(idx, sig, dx, 0.5, 4)
    assert list(e[0]) == [True, False, False, False]
    assert list(e[1]) == [True, True, True, False]
    assert e[2].all() and list(stop) == [True, True, False]


def test_march_row_groups_and_no_tolerance():
    """chunk_size 1: every row is its own group (forward():236-244), so a
    stop never depends on another row; tolerance 0 evaluates everything."""
    idx, sig, dx = _march_case([3, 2, 3], [(0, 0), (2, 1)])
    e, stop, _ = O.march_chunks(idx, sig, dx, 0.5, 1)
    assert list(e[0]) == [True, False, False] and list(e[1]) == [True, True, False]
    assert list(e[2]) == [True, True, False] and list(stop) == [True, False, True]
    e, stop, _ = O.march_chunks(idx, sig, dx, 0.0, 1)
    assert (e == (idx != -1)).all() and not stop.any()


def test_render_view_with_tolerance_and_one_chunk_is_unchanged():
    """A chunk holding every sample ends only after the whole row, so early
    termination changes nothing; a small chunk skips samples behind opaque
    ones and changes the image by less than the tolerance allows."""
    from _cases import shell, shell_rays
    sc = shell()
    rs, rd = shell_rays(width=16, seed=3)
    a = O.render_view(sc, rs, rd, compat=False)
    b = O.render_view(sc, rs, rd, compat=False, tolerance=1e-3, chunk_size=1 << 30)
    for k in ("rgb", "depth", "missed"):
        assert np.array_equal(a[k], b[k])
    c = O.render_view(sc, rs, rd, compat=False, hierarchical=False, tolerance=1e-2, chunk_size=64)
    d = O.render_view(sc, rs, rd, compat=False, hierarchical=False)
    skipped = (c["c_idx"] != -1) & ~c["c_evaluated"]
    assert skipped.sum() > 0 and c["c_stopped"].sum() > 0
    # a stopped ray's remaining transmittance is below the tolerance, and
    # skipped samples could only have added that much
    assert np.abs(c["rgb"] - d["rgb"]).max() <= 1e-2 + 1e-6
