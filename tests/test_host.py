"""Host-side packers on the CPU: the MLP fragments against the matrix-core
operand layout they are read with, and the packed octree boxes."""
import numpy as np

from _cases import shell, shell_octree
from nsvf_amd.field import bf16_bits, bf16_round, pack_layer_bf16, pack_layer_f32
from nsvf_amd.render import node_boxes, pack_tree


def _perm(g, i):
    return 4 * g + i if i < 4 else 16 + 4 * g + (i - 4)


def mfma_16x16x32(a_frag, b_frag):
    """v_mfma_f32_16x16x32_bf16 on one tile: lane 16g + m of A holds row m,
    K-slots 8g..8g+7; lane 16g + n of B holds column n, the same K-slots;
    D[m][n] = sum over the 32 K-slots."""
    D = np.zeros((16, 16))
    for g in range(4):
        for i in range(8):
            D += np.outer(a_frag[16 * g:16 * g + 16, i], b_frag[16 * g:16 * g + 16, i])
    return D


def test_bf16_fragments_compute_w_times_x():
    """A layer packed by pack_layer_bf16 and an input laid out as the gather
    kernel writes it (lane 16g + n, element i = feature 32kt + perm(g, i) of
    sample n) give W @ X through the MFMA's operand semantics."""
    rng = np.random.default_rng(0)
    W = bf16_round(rng.standard_normal((40, 70)).astype(np.float32))   # pads to 48 x 96
    X = bf16_round(rng.standard_normal((70, 16)).astype(np.float32))
    frag = pack_layer_bf16(W, 96)                                        # [3,3,64,8]
    assert frag.shape == (3, 3, 64, 8)
    Xp = np.zeros((96, 16), np.float32)
    Xp[:70] = X
    out = np.zeros((48, 16))
    for mt in range(3):
        for kt in range(3):
            b = np.zeros((64, 8))
            for g in range(4):
                for n in range(16):
                    for i in range(8):
                        b[16 * g + n, i] = Xp[32 * kt + _perm(g, i), n]
            out[16 * mt:16 * mt + 16] += mfma_16x16x32(frag[mt, kt].astype(np.float64), b)
    ref = W.astype(np.float64) @ X.astype(np.float64)
    assert np.abs(out[:40] - ref).max() < 1e-9
    assert not out[40:].any()


def test_bf16_layer_output_feeds_next_layer():
    """The accumulator layout (lane 16g + n holds rows 4g..4g+3 of column n)
    of two consecutive m-tiles, packed as the kernel's epilogue does, is the
    next layer's B operand for one k-tile: features 32kt + perm(g, i)."""
    rng = np.random.default_rng(1)
    Y = rng.standard_normal((32, 16))        # rows of m-tiles 2kt and 2kt + 1
    b = np.zeros((64, 8))
    for g in range(4):
        for n in range(16):
            for j in range(4):
                b[16 * g + n, j] = Y[4 * g + j, n]            # half 0: acc[0][j]
                b[16 * g + n, 4 + j] = Y[16 + 4 * g + j, n]   # half 1: acc[1][j]
    for g in range(4):
        for i in range(8):
            assert np.array_equal(b[16 * g:16 * g + 16, i], Y[_perm(g, i)])


def test_f32_fragments_compute_w_times_x():
    """FP32 fragments (v_mfma_f32_16x16x4_f32: lane 16g + m holds K-slot g):
    k-tile kt = 4t + i, lane 16g + m holds W[16mt + m][16t + 4g + i]."""
    rng = np.random.default_rng(2)
    W = rng.standard_normal((20, 36)).astype(np.float32)    # pads to 32 x 48
    X = rng.standard_normal((36, 16)).astype(np.float32)
    frag = pack_layer_f32(W, 48)                            # [2,12,64]
    Xp = np.zeros((48, 16), np.float32)
    Xp[:36] = X
    out = np.zeros((32, 16))
    for mt in range(2):
        for kt in range(12):
            t, i = divmod(kt, 4)
            for g in range(4):
                a = frag[mt, kt, 16 * g:16 * g + 16].astype(np.float64)
                x = Xp[16 * t + 4 * g + i].astype(np.float64)
                out[16 * mt:16 * mt + 16] += np.outer(a, x)
    assert np.abs(out[:20] - W.astype(np.float64) @ X).max() < 1e-9


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e38, 1e-40], np.float32)
    bits = bf16_bits(x)
    back = (bits.astype(np.uint32) << 16).view(np.float32)
    assert back[0] == 1.0 and back[1] == 1.0         # exactly halfway: to even
    assert back[2] == np.float32(1.0078125)          # above halfway: up
    assert back[3] == -2.5
    assert np.abs(back[:4] - x[:4]).max() <= 2 ** -8


def test_packed_tree_boxes():
    """Leaf boxes are fp32(c - h), fp32(c + h); an internal node's box is the
    exact min/max over its children's boxes, so it contains every descendant
    leaf."""
    sc = shell()
    cen, ch = shell_octree()
    h = np.float32(sc.voxel_size) * np.float32(0.5)
    mn, mx, scale = node_boxes(cen.astype(np.float32), ch.astype(np.int32), h)
    kids = ch[:, :8]
    assert np.array_equal(scale, ch[:, 8])
    leaf = scale == 1
    assert np.array_equal(mn[leaf], (cen[leaf] - h).astype(np.float32))
    assert np.array_equal(mx[leaf], (cen[leaf] + h).astype(np.float32))
    for node in np.nonzero(~leaf)[0][:2000]:
        k = kids[node][kids[node] >= 0]
        assert np.array_equal(mn[node], mn[k].min(0)) and np.array_equal(mx[node], mx[k].max(0))
    root = ch.shape[0] - 1
    assert (mn[root] <= mn[leaf]).all() and (mx[root] >= mx[leaf]).all()


def chain_end(ch, k):
    """Node k, or the end of its chain of one-child nodes."""
    while True:
        kids = ch[k, :8][ch[k, :8] >= 0]
        if len(kids) != 1:
            return k
        k = int(kids[0])


def test_packed_tree_records():
    """Record k, slot s: the (min, scale) and (max, id) of child
    children[k][s], or of the end of its chain of one-child nodes; empty
    slots carry id -1; record T holds the root (or its chain's end) in slot
    0 only."""
    sc = shell()
    cen, ch = shell_octree()
    h = np.float32(sc.voxel_size) * np.float32(0.5)
    mn, mx, scale = node_boxes(cen.astype(np.float32), ch.astype(np.int32), h)
    rec = pack_tree(cen.astype(np.float32), ch.astype(np.int32), h)
    T = ch.shape[0]
    assert rec.shape == (T + 1, 16, 4) and rec.dtype == np.float32
    lo, hi = rec[:, 0::2], rec[:, 1::2]                 # [T+1, 8, 4]
    ids = hi[..., 3].view(np.int32)
    one = ((ch[:, :8] >= 0).sum(1) == 1) & (ch[:, 8] > 1)
    assert one.sum() > T // 4                           # the finest internal level
    want = np.array([[chain_end(ch, c) if c >= 0 else -1 for c in row] for row in ch[:, :8]])
    assert np.array_equal(ids[:T], want)
    assert (scale[want[want >= 0]] == 1).any() and not one[want[want >= 0]].any()
    assert ids[T, 0] == chain_end(ch, T - 1) and (ids[T, 1:] == -1).all()
    # a chain's boxes are all its end's box
    for k in np.nonzero(one)[0][:500]:
        e = chain_end(ch, k)
        assert np.array_equal(mn[k], mn[e]) and np.array_equal(mx[k], mx[e])
    ok = ids >= 0
    k = ids[ok]
    assert np.array_equal(lo[ok][:, :3], mn[k]) and np.array_equal(hi[ok][:, :3], mx[k])
    assert np.array_equal(lo[ok][:, 3].view(np.int32), scale[k])


def test_packed_tree_of_far_apart_voxels():
    """Three voxels far apart hang below the root on chains of one-child
    nodes; the root's packed record holds the three leaves, in the root's
    own child slots, with their exact boxes."""
    from nsvf_amd import geometry, scene_io as IO
    import torch
    pts = np.array([[-3.0, -3.0, -3.0], [0.0, 0.25, 0.0], [3.0, 3.0, 2.75]], np.float32)
    sc = IO.scene_from_points(pts, 0.25, embed_dim=32)
    h = np.float32(sc.voxel_size) * np.float32(0.5)
    cen, ch = geometry.build_easy_octree(torch.from_numpy(sc.points), torch.tensor(h))
    cen, ch = cen.float().numpy(), ch.int().numpy()
    T = ch.shape[0]
    assert T > 3 + 1                                    # the chains exist
    rec = pack_tree(cen, ch, h)
    ids, scale = rec[T - 1, 1::2, 3].view(np.int32), rec[T - 1, 0::2, 3].view(np.int32)
    slots = np.nonzero(ch[T - 1, :8] >= 0)[0]
    assert np.array_equal(np.nonzero(ids >= 0)[0], slots)
    assert (scale[slots] == 1).all()
    assert np.array_equal(ids[slots], [chain_end(ch, c) for c in ch[T - 1, slots]])
    leaf = ids[slots]
    assert np.array_equal(rec[T - 1, 2 * slots, :3], (cen[leaf] - h).astype(np.float32))
    assert np.array_equal(rec[T - 1, 2 * slots + 1, :3], (cen[leaf] + h).astype(np.float32))


def test_split_bf16_pack_from_torch_weights():
    """The training packer (torch ops, here on CPU tensors) equals the
    scene packer's split-BF16 fragments and biases bit for bit: the
    fragment permutation, zero padding, round-to-nearest-even hi and the lo
    residual."""
    import torch
    from nsvf_amd.field import pack_mlp
    from nsvf_amd.scene import FieldMLP
    from nsvf_amd.train import pack_mlp_x3
    rng = np.random.default_rng(7)
    shapes = [(128, 32), (128, 128), (128, 128), (4, 128)]
    W = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    B = [rng.standard_normal(s[0]).astype(np.float32) for s in shapes]
    W[1][3, 5] = np.float32(1.00390625)             # a tie: rounds to even
    host = pack_mlp(FieldMLP(W, B))
    frags, bias = pack_mlp_x3([torch.from_numpy(w) for w in W], [torch.from_numpy(b) for b in B])
    assert torch.equal(frags, host["bf16x3"]) and torch.equal(bias, host["bias"])


def test_split_bf16_training_backward_fragments():
    """pack_mlp_train_x3 against the layout nsvf_mlp_train_backward_x3
    documents, element by element (no GPU: the packer's torch ops on CPU
    tensors): fragments W_0 [8][1], W_1, W_2 [8][4], W_0^T [2][4], W_1^T,
    W_2^T [8][4], each 512 hi then 512 lo; lane l, element i of fragment
    (mt, kt) of A is A[16 mt + l % 16][32 kt + k(i)], k(i) = 4 (l / 16) + i
    for i < 4, 16 + 4 (l / 16) + i - 4 otherwise; hi = bf16(a) (nearest
    even), lo = bf16(a - hi)."""
    import torch
    from nsvf_amd.train import pack_mlp_train_x3
    rng = np.random.default_rng(9)
    shapes = [(128, 32), (128, 128), (128, 128), (4, 128)]
    W = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    got = pack_mlp_train_x3([torch.from_numpy(w) for w in W]).numpy().view(np.uint16)
    mats = [W[0], W[1], W[2], W[0].T, W[1].T, W[2].T]
    want = []
    lane, i = np.arange(64)[:, None], np.arange(8)[None, :]
    kk = np.where(i < 4, 4 * (lane // 16) + i, 16 + 4 * (lane // 16) + i - 4)
    for A in mats:
        A = np.ascontiguousarray(A)
        for mt in range(A.shape[0] // 16):
            for kt in range(A.shape[1] // 32):
                v = A[16 * mt + lane % 16, 32 * kt + kk]            # [64, 8]
                hi = bf16_round(v)
                want += [bf16_bits(hi).reshape(-1), bf16_bits(v - hi).reshape(-1)]
    want = np.concatenate(want)
    assert got.shape == want.shape == (144 * 1024,)
    assert np.array_equal(got, want)
