"""Closed forms and invariants that pin the oracle independently of the
reference build: each check is chosen so that a plausible restatement
mistake (rounding instead of truncating, a dropped tap, a transposed index,
a missing flat-index wrap, a wrong normalisation) fails it."""
import math

import numpy as np

from conftest import smooth_volume

I12 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)


def shift_T(sx, sy, sz):
    T = I12.copy()
    T[9:] = (sx, sy, sz)
    return T


def test_identity_resample_is_copy_except_far_edge(oracle):
    v = smooth_volume((7, 8, 9), 1, background=False)
    out = oracle.resample(v, v.shape, I12, crop=False)
    assert np.array_equal(out[:-1, :-1, :-1], v[:-1, :-1, :-1])
    assert not out[-1].any() and not out[:, -1].any() and not out[:, :, -1].any()
    # the CPU crop shortcut keeps the far edge
    assert np.array_equal(oracle.resample(v, v.shape, I12, crop=True), v)


def test_integer_shift_crop(oracle):
    v = smooth_volume((6, 7, 8), 2, background=False)
    out = oracle.resample(v, v.shape, shift_T(2, -1, 1), crop=True)
    want = np.zeros_like(v)
    want[:-1, 1:, :-2] = v[1:, :-1, 2:]
    assert np.array_equal(out, want)


def test_trilinear_truncates_and_weights(oracle):
    v = np.zeros((3, 3, 4), np.uint8)
    v[:, :, 0::2] = 10
    v[:, :, 1::2] = 21
    assert oracle.sample(v, 0.5, 0.0, 0.0) == 15     # 15.5 truncated, not rounded
    assert oracle.sample(v, 1.25, 1.0, 1.0) == 18    # 21*0.75 + 10*0.25 = 18.25
    ramp = np.tile(np.arange(4, dtype=np.uint8) * 40, (3, 3, 1))
    assert oracle.sample(ramp, 2.5, 0.5, 0.5) == 100
    assert oracle.sample(ramp, 2.999, 0.0, 0.0) == 119  # 119.96 truncated
    assert oracle.sample(ramp, 3.0, 0.0, 0.0) == 0      # ix + 1 >= w: outside
    assert oracle.sample(ramp, -0.001, 0.0, 0.0) == 0


def test_y_weight_order(oracle):
    # taps differ per corner: a transposed weight (y for z) changes the value
    v = np.zeros((2, 2, 2), np.uint8)
    v[0, 1, 0] = 200   # (x=0, y=1, z=0)
    assert oracle.sample(v, 0.0, 0.25, 0.0) == 50
    assert oracle.sample(v, 0.0, 0.0, 0.25) == 0


def test_correlation_closed_forms(oracle):
    x = smooth_volume((10, 11, 12), 3, background=False)
    x = (x // 2).astype(np.uint8)
    y = (2 * x.astype(np.int32) + 1).astype(np.uint8)         # exact affine intensity map
    s = oracle.corr_sums(x, y)
    assert oracle.corr_cost_from_sums(s, x.size) == 0.0
    const = np.full_like(x, 7)
    assert oracle.corr_cost_from_sums(oracle.corr_sums(x, const), x.size) == 1.0
    # sums by brute force
    xs, ys = x.astype(np.int64), y.astype(np.int64)
    assert list(s) == [xs.sum(), (xs * xs).sum(), ys.sum(), (ys * ys).sum(), (xs * ys).sum()]


def test_mi_identity_closed_form(oracle):
    v = smooth_volume((9, 10, 11), 4)   # zero background margin: the far edge samples 0 == value
    h = oracle.mi_hist(v, v, I12)
    counts = np.bincount(v.ravel(), minlength=256)
    assert np.array_equal(h, np.diag(counts))
    nz = counts[counts > 0].astype(np.float64)
    want = float((nz * np.log(nz)).sum() / 65536.0)    # -sum h log(h/(h*h)) / bins^2
    got = oracle.mi_cost(v, v, I12)
    assert abs(got - want) <= 1e-12 * abs(want)
    # a one-to-one relabelling of the moving image leaves MI unchanged
    perm = np.concatenate([[0], 1 + np.random.default_rng(5).permutation(255)]).astype(np.uint8)
    assert abs(oracle.mi_cost(v, perm[v], I12) - got) <= 1e-12 * abs(got)


def test_mi_row_sums_are_from_histogram(oracle):
    A, B = smooth_volume((9, 12, 10), 6), smooth_volume((11, 8, 13), 7)
    T = np.array([0.9, 0.1, 0.05, -0.1, 1.1, 0.0, 0.02, 0.0, 0.95, 1.3, -0.7, 2.1], np.float32)
    for bins in (256, 64):
        h = oracle.mi_hist(A, B, T, bins)
        assert h.sum() == A.size
        assert np.array_equal(h.sum(1), oracle.value_hist(A, bins))


def test_mi64_block_sums_of_mi256(oracle):
    """The 64-bin joint histogram (both values >> 2) is the 4x4 block sum of
    the 256-bin one, which is pinned against the reference itself; a column
    or row binning slip (t >> 3, swapped axes) breaks the identity."""
    A, B = smooth_volume((13, 11, 12), 16), smooth_volume((10, 14, 9), 17)
    rng = np.random.default_rng(18)
    Ts = [I12, np.array([0.9, 0.1, 0.05, -0.1, 1.1, 0.0, 0.02, 0.0, 0.95, 1.3, -0.7, 2.1],
                        np.float32),
          np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 40, 0, 0], np.float32)]   # all outside
    for _ in range(4):
        T = I12.copy()
        T[:9] += rng.uniform(-0.15, 0.15, 9).astype(np.float32)
        T[9:] = rng.uniform(-2, 2, 3).astype(np.float32)
        Ts.append(T)
    for T in Ts:
        h256 = oracle.mi_hist(A, B, T, 256).astype(np.int64)
        h64 = oracle.mi_hist(A, B, T, 64)
        assert np.array_equal(h64, h256.reshape(64, 4, 64, 4).sum(axis=(1, 3)))
    assert np.array_equal(oracle.value_hist(A, 64), oracle.value_hist(A, 256).reshape(64, 4).sum(1))


def test_mi64_relabelling_closed_form(oracle):
    """Identity transform at 64 bins: a diagonal histogram of v >> 2 and
    MI = sum h log h / 64^2; a one-to-one relabelling of the 64 labels of the
    moving image leaves the cost unchanged."""
    v = smooth_volume((9, 10, 11), 19)
    h = oracle.mi_hist(v, v, I12, 64)
    counts = np.bincount((v >> 2).ravel(), minlength=64)
    assert np.array_equal(h, np.diag(counts))
    nz = counts[counts > 0].astype(np.float64)
    want = float((nz * np.log(nz)).sum() / 4096.0)
    got = oracle.mi_cost(v, v, I12, 64)
    assert abs(got - want) <= 1e-12 * abs(want)
    perm = np.concatenate([[0], 1 + np.random.default_rng(20).permutation(63)])
    moved = (perm[v >> 2] * 4 + (v & 3)).astype(np.uint8)
    assert abs(oracle.mi_cost(v, moved, I12, 64) - got) <= 1e-12 * abs(got)


def test_quantise_limits(oracle):
    v = np.linspace(-3, 5, 1000, dtype=np.float32).reshape(10, 10, 10)
    q = oracle.normalize(v, 255.99)
    assert q.min() == 0 and q.max() == 255
    assert np.all(np.diff(q.ravel().astype(int)) >= 0)


def test_poisson_flat_index_wrap(oracle):
    # a delta at the end of the first row of a 3x3x3 grid reaches the start of
    # the second row through the flat-index +1 neighbour
    nd = np.zeros((3, 3, 3, 3), np.float32)
    nd[0, 0, 2, 0] = 6.0
    out = oracle.poisson(nd, 1)
    k = np.float32(-0.5) / np.float32(3.0)
    s0 = np.float32(6.0) * k
    nb = k * (np.float32(0.0) - s0)
    assert out[0, 1, 0, 0] == nb          # flat p+1: wrapped neighbour
    assert out[0, 0, 1, 0] == nb          # p-1
    assert out[0, 1, 2, 0] == nb          # p+w
    assert out[1, 0, 2, 0] == nb          # p+wh
    assert out[0, 0, 2, 0] == k * np.float32(6.0)
    assert out[0, 2, 2, 0] == 0.0 and out[0, 1, 1, 0] == 0.0
    assert not out[..., 1:].any()
    # linear in the right-hand side (exact for power-of-two scaling)
    rng = np.random.default_rng(8)
    a = rng.standard_normal((5, 6, 7, 3)).astype(np.float32)
    assert np.array_equal(oracle.poisson(a * 4, 12), oracle.poisson(a, 12) * 4)


def test_cdm_identities(oracle):
    Is = smooth_volume((8, 9, 10), 9, background=False)
    zero = np.zeros(Is.shape + (3,), np.float32)
    assert np.array_equal(oracle.compose(Is, zero), Is)
    cur = np.random.default_rng(10).standard_normal(Is.shape + (3,)).astype(np.float32)
    assert np.array_equal(oracle.accumulate(cur, zero), cur)
    dd = np.random.default_rng(11).standard_normal(Is.shape + (3,)).astype(np.float32)
    assert np.array_equal(oracle.smooth(cur, dd, 0.0), dd)          # smoothing 0: untouched
    c = np.full(Is.shape + (3,), 1.5, np.float32)
    sm = oracle.smooth(c, dd, 0.5)
    assert np.allclose(sm[2:-2, 2:-2, 2:-2], 1.5)                  # interior of a constant
    # 'd' gradient of identical images is zero with zero cost
    nd, cost = oracle.gradient(Is, Is, "d")
    assert not nd.any() and cost == 0.0


def test_pyramid_closed_forms(oracle):
    v = np.full((6, 7, 8), 77, np.uint8)
    assert np.all(oracle.downsample(v) == 77)
    b = np.arange(8, dtype=np.uint8).reshape(2, 2, 2) * 30        # sum 840 -> 105
    assert oracle.downsample(b)[0, 0, 0] == 840 >> 3
    c = np.full((4, 5, 6, 3), 2.0, np.float32)
    up = oracle.upsample(c, (8, 10, 12))
    assert np.all(up[:-2, :-2, :-2] == 2.0)
    assert not up[-2:].any() and not up[:, -2:].any() and not up[:, :, -2:].any()


def test_cdm_improved_trend(oracle):
    assert oracle.improved([5, 4, 3, 2, 1, 0.5])[0]
    assert not oracle.improved([1, 2, 3, 4, 5, 6])[0]
    ok, c = oracle.improved([9, 8, 7, 6, 5, 4, 3, 2])
    assert ok and len(c) == 7 and c[0] == 8
    assert oracle.improved([3, 1])[0]


def test_affine_matrix_identity_and_shift(oracle):
    d = np.array([20, 30, 40], np.int32)
    T = oracle.affine_matrix([0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], d, (1, 1, 1), d, (1, 1, 1))
    assert np.array_equal(T, I12)
    T = oracle.affine_matrix([2, -3, 1.5, 0, 0, 0, 1, 1, 1, 0, 0, 0], d, (1, 1, 1), d, (1, 1, 1))
    assert np.allclose(T, shift_T(2, -3, 1.5))
    # a rotation about z by pi/2 maps x to y (reference's left-handed Euler)
    T = oracle.affine_matrix([0, 0, 0, 0, 0, math.pi / 2, 1, 1, 1, 0, 0, 0], d, (1, 1, 1), d, (1, 1, 1))
    assert np.allclose(T[:9].reshape(3, 3), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]], atol=1e-6)


def test_gaussian_closed_forms(oracle):
    """filter::gaussian: constants are fixed points, an interior impulse
    spreads as the outer product of binomial rows C(2b, k) / 4^b, mass is
    kept away from the edges, variance is b/2 per axis, and the three
    channels of a field are filtered independently."""
    from math import comb
    const = np.full((7, 8, 9), 37.25, np.float32)
    for b in (1, 2, 3):
        assert np.array_equal(oracle.gaussian(const, b), const)
    n = 15
    imp = np.zeros((n, n, n), np.float32)
    imp[7, 7, 7] = 4096.0
    for b in (1, 2, 3):
        out = oracle.gaussian(imp, b).astype(np.float64)
        row = np.zeros(n)
        row[7 - b:8 + b] = [comb(2 * b, k) / 4.0 ** b for k in range(2 * b + 1)]
        expect = 4096.0 * np.einsum("i,j,k->ijk", row, row, row)
        assert np.allclose(out, expect, rtol=1e-6, atol=1e-9)
        assert abs(out.sum() - 4096.0) < 1e-3
        pos = np.arange(n) - 7.0
        var_x = (out.sum((0, 1)) * pos ** 2).sum() / out.sum()
        assert abs(var_x - b / 2.0) < 1e-6
    rng = np.random.default_rng(3)
    f = rng.standard_normal((6, 5, 4, 3)).astype(np.float32)
    g = oracle.gaussian(f, 2)
    for c in range(3):
        assert np.array_equal(g[..., c], oracle.gaussian(np.ascontiguousarray(f[..., c]), 2))
    # replicate edges: a ramp along x keeps its interior slope and flattens at the ends
    ramp = np.broadcast_to(np.arange(9, dtype=np.float32), (3, 4, 9)).copy()
    r = oracle.gaussian(ramp, 1)[1, 1]
    assert np.allclose(r[1:-1], np.arange(1, 8)) and r[0] == 0.25 and r[-1] == 7.75


def _linear_field(shape, A):
    """d(x) = A x over the (z, y, x) grid, x = (i, j, k) voxel coordinates."""
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    X = np.stack([x, y, z], -1)
    return (X @ np.asarray(A, np.float64).T).astype(np.float32)


def test_jacobian_closed_forms(oracle):
    """jacobian_determinant_dis of d(x) = A x is det(I + 2A) (central
    differences not halved) with 1 on the edges; jacobian_determinant of a
    mapping m(x) = B x is 8 det(B) with the reference's 0.125 applied to the
    third cofactor term only.  A and B are multiples of 1/8, so every float
    and double step is exact."""
    A = np.array([[0.125, -0.25, 0.0], [0.375, 0.0, 0.125], [-0.125, 0.25, 0.5]])
    shape = (6, 7, 8)
    J = oracle.jacobian_determinant_dis(_linear_field(shape, A))
    inner = J[1:-1, 1:-1, 1:-1]
    assert np.all(inner == np.float32(np.linalg.det(np.eye(3) + 2 * A)))
    edge = np.ones(shape, bool)
    edge[1:-1, 1:-1, 1:-1] = False
    assert np.all(J[edge] == 1.0)
    B = np.array([[1.0, 0.5, 0.25], [0.125, 1.25, -0.5], [0.75, 0.25, 2.0]])
    Jm = oracle.jacobian_determinant(_linear_field(shape, B))
    full = 8 * (B[0, 0] * (B[1, 1] * B[2, 2] - B[2, 1] * B[1, 2])
                + B[1, 0] * (B[2, 1] * B[0, 2] - B[0, 1] * B[2, 2]))
    quirk = 8 * B[2, 0] * (B[0, 1] * B[1, 2] - B[1, 1] * B[0, 2]) * 0.125
    assert np.all(Jm[1:-1, 1:-1, 1:-1] == np.float32(full + quirk))
    assert np.all(Jm[edge] == 0.0)
    assert np.float32(full + quirk) != np.float32(8 * np.linalg.det(B))


def test_mapping_identities(oracle):
    """A zero field maps to the voxel grid; T is applied after adding the
    index; an integer-shift T with a zero inverse field inverts to x - shift;
    compose_mapping through the identity mapping copies the image except at
    the far edges (where the trilinear cell is incomplete)."""
    shape = (5, 6, 7)
    zero = np.zeros(shape + (3,), np.float32)
    grid = oracle.displacement_to_mapping(zero)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in shape], indexing="ij")
    assert np.array_equal(grid, np.stack([x, y, z], -1))
    d = np.full(shape + (3,), 0.5, np.float32)
    m = oracle.displacement_to_mapping(d, shift_T(2, -1, 3))
    assert np.array_equal(m, np.stack([x + 2.5, y - 0.5, z + 3.5], -1))
    assert np.array_equal(oracle.mapping_to_displacement(m), np.stack([x * 0 + 2.5, y * 0 - 0.5,
                                                                         z * 0 + 3.5], -1))
    inv = oracle.inv_displacement_to_mapping(zero, shift_T(2, -1, 3), (4, 5, 6))
    z2, y2, x2 = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in (4, 5, 6)], indexing="ij")
    assert np.array_equal(inv, np.stack([x2 - 2, y2 + 1, z2 - 3], -1))
    img = (np.arange(np.prod(shape)) % 251).astype(np.uint8).reshape(shape)
    out = oracle.compose_mapping(img, grid)
    assert np.array_equal(out[:-1, :-1, :-1], img[:-1, :-1, :-1])
    assert not out[-1].any() and not out[:, -1].any() and not out[:, :, -1].any()


def test_masked_correlation_closed_forms(oracle):
    """r = 1 for an exact increasing affine relation on the masked voxels even
    when the unmasked voxels break it; r = -1 for a decreasing one; the mask
    is y > 0 only."""
    rng = np.random.default_rng(2)
    y = rng.integers(1, 100, 500).astype(np.uint8)
    y[::7] = 0
    x = (2 * y.astype(np.int32) + 3).astype(np.uint8)
    x[y == 0] = rng.integers(0, 255, int((y == 0).sum()))
    assert abs(oracle.correlation(x, y, True) - 1.0) < 1e-12
    assert oracle.correlation(x, y, False) < 0.99
    xd = (250 - 2 * y.astype(np.int32)).astype(np.uint8)
    assert abs(oracle.correlation(xd, y, True) + 1.0) < 1e-12
    ref_r = np.corrcoef(x[y > 0].astype(np.float64), y[y > 0].astype(np.float64))[0, 1]
    assert abs(oracle.correlation(x, y, True) - ref_r) < 1e-12


def test_match_fov_shift(oracle):
    """An integer shift of +2 along x moves the last two template columns
    outside a subject grid of the same shape; nothing else is zeroed."""
    It = np.full((4, 5, 6), 9, np.uint8)
    out = oracle.match_fov(It, shift_T(2, 0, 0), (4, 5, 6))
    assert not out[:, :, 4:].any() and np.all(out[:, :, :4] == 9)
    assert np.array_equal(oracle.match_fov(It, I12, (4, 5, 6)), It)
