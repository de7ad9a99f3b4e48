"""GPU parity of the affine path: resample, correlation sums/cost, MI
histogram/cost, quantisation and pyramid -- HIP kernels through the C ABI
against the oracle on the same seeded inputs.

Bar: bit-exact for every integer / byte result (resampled voxels, int64
correlation sums, int32 joint histograms, quantised volumes); costs within
rel. 1e-9 of the oracle's FP64 value (the spec allows 1e-5).
"""
import numpy as np
import pytest

from conftest import random_params, smooth_volume

pytestmark = pytest.mark.gpu

COST_RTOL = 1e-9


def _T(oracle, p, fshape, tshape, fvs=(1, 1, 1), tvs=(1, 1, 1)):
    return oracle.affine_matrix(p, np.array(fshape[::-1], np.int32), fvs,
                                np.array(tshape[::-1], np.int32), tvs)


SHAPES = [((20, 24, 28), (20, 24, 28)), ((17, 23, 31), (19, 21, 29)), ((9, 5, 70), (9, 5, 70))]


@pytest.mark.parametrize("fshape,tshape", SHAPES)
def test_resample_bit_exact(ctx, oracle, fshape, tshape):
    rng = np.random.default_rng(1)
    src = smooth_volume(tshape, 2)
    vol = ctx.volume(src)
    for p in random_params(rng, 6):
        T = _T(oracle, p, fshape, tshape)
        for crop in (True, False):
            got = ctx.resample(vol, fshape, T, crop=crop)
            want = oracle.resample(src, fshape, T, crop=crop)
            assert np.array_equal(got, want)


def test_resample_integer_shift_crop(ctx, oracle):
    src = smooth_volume((12, 14, 16), 3, background=False)
    vol = ctx.volume(src)
    for sh in ([0, 0, 0], [2, -1, 3], [-3, 5, 1], [20, 0, 0]):
        T = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, *sh], np.float32)
        got_c = ctx.resample(vol, src.shape, T, crop=True)
        assert np.array_equal(got_c, oracle.resample(src, src.shape, T, crop=True))
        got_n = ctx.resample(vol, src.shape, T, crop=False)
        assert np.array_equal(got_n, oracle.resample(src, src.shape, T, crop=False))
    # identity: the crop keeps the far-edge slab, interpolation zeroes it
    T = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
    assert np.array_equal(ctx.resample(vol, src.shape, T, crop=True), src)
    nc = ctx.resample(vol, src.shape, T, crop=False)
    assert np.array_equal(nc[:-1, :-1, :-1], src[:-1, :-1, :-1])
    assert not nc[-1].any() and not nc[:, -1].any() and not nc[:, :, -1].any()


def test_resample_nan_and_far_transforms(ctx, oracle):
    src = smooth_volume((10, 11, 12), 4)
    vol = ctx.volume(src)
    for T in (np.array([np.nan] * 12, np.float32),
              np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 1e9, 0, 0], np.float32),
              np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, -1e9, 0, 0], np.float32)):
        got = ctx.resample(vol, src.shape, T, crop=False)
        assert np.array_equal(got, oracle.resample(src, src.shape, T, crop=False))
        assert not got.any()


@pytest.mark.parametrize("fshape,tshape", SHAPES)
def test_corr_sums_bit_exact(ctx, oracle, fshape, tshape):
    rng = np.random.default_rng(5)
    A, B = smooth_volume(fshape, 6), smooth_volume(tshape, 7)
    va, vb = ctx.volume(A), ctx.volume(B)
    P = random_params(rng, 16)
    Ts = np.stack([_T(oracle, p, fshape, tshape) for p in P])
    got = ctx.corr_sums(va, vb, Ts)
    for k, T in enumerate(Ts):
        r = oracle.resample(B, fshape, T, crop=False)
        assert np.array_equal(got[k], oracle.corr_sums(A, r))


@pytest.mark.parametrize("fshape,tshape", SHAPES + [((30, 30, 30), (20, 20, 20))])
def test_corr_cost_batch(ctx, oracle, fshape, tshape):
    rng = np.random.default_rng(8)
    A, B = smooth_volume(fshape, 9), smooth_volume(tshape, 10)
    va, vb = ctx.volume(A), ctx.volume(B)
    P = random_params(rng, 20)
    Ts = np.stack([_T(oracle, p, fshape, tshape) for p in P])
    # integer-shift candidates exercise the crop shortcut
    Ts[0] = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    Ts[1] = [1, 0, 0, 0, 1, 0, 0, 0, 1, 2, -1, 1]
    got = ctx.cost_batch(va, vb, Ts, cost="corr")
    want = np.array([oracle.corr_cost(A, B, T) for T in Ts])
    np.testing.assert_allclose(got, want, rtol=COST_RTOL, atol=1e-12)


@pytest.mark.parametrize("bins", [256, 64])
@pytest.mark.parametrize("fshape,tshape", SHAPES)
def test_mi_hist_bit_exact(ctx, oracle, fshape, tshape, bins):
    rng = np.random.default_rng(11)
    A, B = smooth_volume(fshape, 12), smooth_volume(tshape, 13)
    va, vb = ctx.volume(A), ctx.volume(B)
    P = random_params(rng, 8)
    Ts = np.stack([_T(oracle, p, fshape, tshape) for p in P])
    got = ctx.mi_hist(va, vb, Ts, bins=bins)
    for k, T in enumerate(Ts):
        assert np.array_equal(got[k], oracle.mi_hist(A, B, T, bins=bins))


@pytest.mark.parametrize("bins", [256, 64])
def test_mi_cost_batch(ctx, oracle, bins):
    rng = np.random.default_rng(14)
    A, B = smooth_volume((22, 26, 30), 15), smooth_volume((22, 26, 30), 16)
    va, vb = ctx.volume(A), ctx.volume(B)
    P = random_params(rng, 33)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape) for p in P])
    got = ctx.cost_batch(va, vb, Ts, cost="mi", bins=bins)
    want = np.array([oracle.mi_cost(A, B, T, bins=bins) for T in Ts])
    np.testing.assert_allclose(got, want, rtol=COST_RTOL)


def test_mi_cost_sequence_reuses_cleared_histograms(ctx, oracle):
    """The entropy kernels leave the histogram buffer zero for the next
    launch; launches of growing and shrinking size, both bin counts, and the
    raw histogram and autotune entry points (which do not clear it) in
    between must not change any cost."""
    rng = np.random.default_rng(24)
    A, B = smooth_volume((18, 22, 26), 25), smooth_volume((20, 20, 24), 26)
    va, vb = ctx.volume(A), ctx.volume(B)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, 40)])
    want = {bins: np.array([oracle.mi_cost(A, B, T, bins=bins) for T in Ts]) for bins in (256, 64)}
    steps = [(256, 3), (256, 40), (64, 7), (256, 1), "hist", (256, 12), (64, 40), "tune",
             (64, 5), (256, 40), (256, 2)]
    for step in steps:
        if step == "hist":
            ctx.mi_hist(va, vb, Ts[:9], bins=256)
            continue
        if step == "tune":
            ctx.autotune(va, vb, Ts[:6], "mi", 256, reps=1)
            continue
        bins, k = step
        got = ctx.cost_batch(va, vb, Ts[:k], cost="mi", bins=bins)
        np.testing.assert_allclose(got, want[bins][:k], rtol=COST_RTOL)
        again = ctx.cost_batch(va, vb, Ts[:k], cost="mi", bins=bins)
        assert np.array_equal(got, again)


def test_large_launch_z_table_is_bit_identical(ctx, oracle):
    """Launches of at least 2^25 samples read the z terms of the transformed
    positions from a table built by a separate kernel; smaller ones compute
    them inline.  130 candidates on a 64^3 pair (34M samples) take the table
    path, batches of 10 do not: costs and histograms must agree bit for bit,
    and with the oracle."""
    rng = np.random.default_rng(27)
    A, B = smooth_volume((64, 64, 64), 28), smooth_volume((64, 64, 64), 29)
    va, vb = ctx.volume(A), ctx.volume(B)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, 130)])
    assert 130 * A.size >= 1 << 25 and 10 * A.size < 1 << 25
    for cost, bins in (("mi", 64), ("mi", 256), ("corr", 256)):
        big = ctx.cost_batch(va, vb, Ts, cost=cost, bins=bins)
        small = np.concatenate([ctx.cost_batch(va, vb, Ts[i:i + 10], cost=cost, bins=bins)
                                for i in range(0, len(Ts), 10)])
        assert np.array_equal(big.view(np.uint64), small.view(np.uint64)), (cost, bins)
    h = ctx.mi_hist(va, vb, Ts, bins=64)
    for k in (0, 77, 129):
        assert np.array_equal(h[k], oracle.mi_hist(A, B, Ts[k], bins=64))
    s = ctx.corr_sums(va, vb, Ts)
    for k in (3, 128):
        assert np.array_equal(s[k], oracle.corr_sums(A, oracle.resample(B, A.shape, Ts[k], crop=False)))


@pytest.mark.parametrize("shape,K", [((40, 44, 36), 7), ((64, 64, 64), 130), ((23, 17, 31), 3)])
@pytest.mark.parametrize("bins", [256, 64])
def test_launch_shapes_are_identical(ctx, oracle, shape, K, bins):
    """Histograms and costs do not depend on the launch shape (planes per
    work item, workgroups per CU, pinned with set_launch_shape): bit for bit
    across shapes and against the oracle; 130 candidates of 64^3 take the
    z-table kernels."""
    rng = np.random.default_rng(41)
    A, B = smooth_volume(shape, 42), smooth_volume(shape[::-1], 43)
    va, vb = ctx.volume(A), ctx.volume(B)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, K)])
    ref_h = ref_c = None
    for zc, per_cu in ((16, 2), (8, 0), (32, 4), (4, 16)):
        ctx.set_launch_shape(va, vb, K, "mi", bins, zc=zc, per_cu=per_cu)
        h = ctx.mi_hist(va, vb, Ts, bins=bins)
        c = ctx.cost_batch(va, vb, Ts, "mi", bins)
        if ref_h is None:
            ref_h, ref_c = h, c
            for k in (0, K - 1):
                assert np.array_equal(h[k], oracle.mi_hist(A, B, Ts[k], bins=bins))
        assert np.array_equal(h, ref_h), (zc, per_cu)
        assert np.array_equal(c.view(np.uint64), ref_c.view(np.uint64)), (zc, per_cu)
    import tipl_amd
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.set_launch_shape(va, vb, K, "mi", 100)


def test_batch_equals_single(ctx, oracle):
    rng = np.random.default_rng(17)
    A, B = smooth_volume((24, 20, 28), 18), smooth_volume((24, 20, 28), 19)
    va, vb = ctx.volume(A), ctx.volume(B)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, 12)])
    for cost in ("mi", "corr"):
        batch = ctx.cost_batch(va, vb, Ts, cost=cost)
        single = np.array([ctx.cost_batch(va, vb, T[None], cost=cost)[0] for T in Ts])
        assert np.array_equal(batch, single)


def test_cost_batch_params_matches_matrices(ctx, oracle):
    rng = np.random.default_rng(20)
    A, B = smooth_volume((16, 18, 20), 21), smooth_volume((18, 16, 22), 22)
    va, vb = ctx.volume(A), ctx.volume(B)
    P = random_params(rng, 10)
    fvs, tvs = (1.0, 1.2, 0.9), (1.1, 1.0, 1.3)
    Ts = np.stack([_T(oracle, p, A.shape, B.shape, fvs, tvs) for p in P])
    a = ctx.cost_batch_params(va, fvs, vb, tvs, P, cost="mi")
    b = ctx.cost_batch(va, vb, Ts, cost="mi")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.float32, np.int16])
def test_quantise_on_upload(ctx, oracle, dtype):
    rng = np.random.default_rng(23)
    v = (rng.normal(0, 300, (14, 15, 16))).astype(dtype)
    vol = ctx.volume(v)
    got = vol.read()
    assert np.array_equal(got, oracle.normalize(v, 255.99))
    s, s2, h = vol.stats()
    assert s == int(got.astype(np.int64).sum())
    assert s2 == int((got.astype(np.int64) ** 2).sum())
    assert np.array_equal(h, np.bincount(got.ravel(), minlength=256))


def test_constant_volume_quantises_to_zero(ctx):
    vol = ctx.volume(np.full((5, 6, 7), 3.5, np.float32))
    assert not vol.read().any()


@pytest.mark.parametrize("shape", [(9, 10, 11), (1, 1, 1), (2, 3, 4), (16, 16, 16)])
def test_downsample_bit_exact(ctx, oracle, shape):
    v = smooth_volume(shape, 24, background=False) if min(shape) > 2 else \
        np.random.default_rng(0).integers(0, 256, shape).astype(np.uint8)
    d = ctx.volume(v).downsample()
    assert d.shape == tuple((s + 1) // 2 for s in shape)
    assert np.array_equal(d.read(), oracle.downsample(v))


def test_tiny_and_degenerate_volumes(ctx, oracle):
    # 1-voxel and 2-voxel-thick volumes: every sample is outside or at an edge
    for shape in [(1, 1, 1), (2, 2, 2), (1, 5, 7)]:
        A = np.random.default_rng(1).integers(0, 256, shape).astype(np.uint8)
        va = ctx.volume(A)
        T = _T(oracle, [0.3, 0.1, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], shape, shape)
        assert np.array_equal(ctx.mi_hist(va, va, T[None])[0], oracle.mi_hist(A, A, T))
        assert np.array_equal(ctx.resample(va, shape, T, crop=False),
                              oracle.resample(A, shape, T, crop=False))
        got = ctx.cost_batch(va, va, T[None], cost="corr")[0]
        assert got == oracle.corr_cost(A, A, T)


def test_invalid_arguments_raise(ctx):
    import tipl_amd
    v = ctx.volume(np.zeros((4, 4, 4), np.uint8))
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.cost_batch(v, v, np.zeros((0, 12), np.float32))
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.mi_hist(v, v, np.zeros((1, 12), np.float32), bins=100)
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.volume(np.zeros((0, 4, 4), np.uint8))


# ---- full-size properties at the 182x218x182 workload ----------------------

@pytest.fixture(scope="module")
def tissue_u8():
    from tipl_amd import phantoms
    f, m, truth = phantoms.tissue_pair(1000)
    return phantoms.quantize(f), phantoms.quantize(m), truth


def test_full_size_mi_properties(ctx, oracle, tissue_u8):
    F, M, truth = tissue_u8
    vf, vm = ctx.volume(F), ctx.volume(M)
    rng = np.random.default_rng(1001)
    P = np.concatenate([np.array([[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0]], np.float32),
                        random_params(rng, 7, t=8, r=0.17, s=0.05, a=0.03)])
    Ts = np.stack([_T(oracle, p, F.shape, M.shape) for p in P])
    H = ctx.mi_hist(vf, vf, Ts[:1])[0]
    # identity on the same volume: diagonal, except far-edge voxels (sampled 0)
    off = H - np.diag(np.diag(H))
    assert off[:, 1:].sum() == 0
    fh = np.bincount(F.ravel(), minlength=256)
    H = ctx.mi_hist(vf, vm, Ts)
    for k in range(len(Ts)):
        assert H[k].sum() == F.size                       # every voxel counted once
        assert np.array_equal(H[k].sum(1), fh)            # rows = from histogram
    for k in (0, 3):
        assert np.array_equal(H[k], oracle.mi_hist(F, M, Ts[k]))
    costs = ctx.cost_batch(vf, vm, Ts, cost="mi")
    for k in (0, 5):
        assert abs(costs[k] - oracle.mi_cost(F, M, Ts[k])) <= COST_RTOL * abs(costs[k])


def test_full_size_corr_sums(ctx, oracle, tissue_u8):
    F, M, _ = tissue_u8
    vf, vm = ctx.volume(F), ctx.volume(M)
    rng = np.random.default_rng(1002)
    Ts = np.stack([_T(oracle, p, F.shape, M.shape) for p in random_params(rng, 3, t=8, r=0.17)])
    got = ctx.corr_sums(vf, vm, Ts)
    for k, T in enumerate(Ts):
        r = oracle.resample(M, F.shape, T, crop=False)
        assert np.array_equal(got[k], oracle.corr_sums(F, r))


def _inside_params(K, scale, seed):
    rng = np.random.default_rng(seed)
    P = random_params(rng, K, t=1.0, r=0.05, s=0.02, a=0.01)
    P[:, 6:9] *= scale
    return P


@pytest.mark.parametrize("tshape", [(60, 64, 60), (256, 256, 256), (257, 256, 256)])
def test_inside_columns_bit_exact(ctx, oracle, tshape):
    """Correlation waves whose columns all map inside the `to` volume skip the
    per-sample bounds test and form the cell index in FP32 (up to 2^24
    voxels); the 256^3 case puts cell indices near 2^24, and 257x256x256
    (over 2^24) takes the checked path.  Correlation sums, and the joint
    histograms (256 and 64 bins) of the same transforms, bit-exact against
    the oracle, for transforms with every column inside and for transforms
    partly outside."""
    fshape = (40, 48, 40) if tshape[0] == 60 else (64, 64, 64)
    scale = 1.0 if tshape[0] == 60 else 3.9
    A = smooth_volume(fshape, 21)
    B = (smooth_volume(tshape, 22) if tshape[0] == 60 else
         np.random.default_rng(23).integers(0, 256, tshape, dtype=np.uint8))
    va, vb = ctx.volume(A), ctx.volume(B)
    P = _inside_params(6, scale, 24)
    P[4, 0:3] = [25.0, -30.0, 12.0]          # partly outside
    P[5, 0:3] = [0.0, 0.0, 200.0]            # mostly outside
    Ts = np.stack([_T(oracle, p, fshape, tshape) for p in P])
    for bins in (256, 64):
        got = ctx.mi_hist(va, vb, Ts, bins=bins)
        for k, T in enumerate(Ts):
            assert np.array_equal(got[k], oracle.mi_hist(A, B, T, bins=bins)), (bins, k)
    got = ctx.corr_sums(va, vb, Ts)
    for k, T in enumerate(Ts):
        r = oracle.resample(B, fshape, T, crop=False)
        assert np.array_equal(got[k], oracle.corr_sums(A, r)), k
    vb.free()
    va.free()
