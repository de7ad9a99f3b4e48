"""Pins the oracle (oracle/oracle.c) to the reference: every restated
function against the reference headers compiled by oracle/Makefile into
oracle/_ref/libtiplref.so, on seeded inputs, bit for bit.  Skipped where the
reference build is not present (its results are also frozen in
tests/golden/, see test_oracle_golden.py)."""
import numpy as np
import pytest

from conftest import prep_images, random_params, smooth_volume


def _T(oracle, p, fshape, tshape, fvs=(1, 1, 1), tvs=(1, 1, 1)):
    return oracle.affine_matrix(p, np.array(fshape[::-1], np.int32), fvs,
                                np.array(tshape[::-1], np.int32), tvs)


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def test_affine_matrix_and_inverse(oracle, ref):
    rng = np.random.default_rng(0)
    for _ in range(100):
        p = random_params(rng, 1, t=20, r=1.2, s=0.4, a=0.2)[0]
        fd = rng.integers(5, 300, 3).astype(np.int32)
        td = rng.integers(5, 300, 3).astype(np.int32)
        fvs = rng.uniform(0.5, 3, 3).astype(np.float32)
        tvs = rng.choice([1.0, 0.7, 2.0], 3).astype(np.float32)
        a = oracle.affine_matrix(p, fd, fvs, td, tvs)
        b = ref.affine_matrix(p, fd, fvs, td, tvs)
        assert np.array_equal(_bits(a), _bits(b))
        ia, oka = oracle.matrix_inverse(a)
        ib, okb = ref.matrix_inverse(a)
        assert oka == okb and np.array_equal(_bits(ia), _bits(ib))


SHAPES = [((14, 18, 22), (14, 18, 22)), ((11, 19, 15), (13, 17, 21)), ((20, 20, 20), (12, 12, 12))]


@pytest.mark.parametrize("fshape,tshape", SHAPES)
def test_resample_and_costs(oracle, ref, fshape, tshape):
    rng = np.random.default_rng(1)
    A, B = smooth_volume(fshape, 2), smooth_volume(tshape, 3)
    P = random_params(rng, 12)
    Ts = [_T(oracle, p, fshape, tshape) for p in P]
    Ts.append(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 1, -2, 0], np.float32))   # crop shortcut
    Ts.append(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32))
    for T in Ts:
        assert np.array_equal(oracle.resample(B, fshape, T), ref.resample(B, fshape, T))
        assert oracle.corr_cost(A, B, T) == ref.corr_cost(A, B, T)
        assert np.array_equal(oracle.mi_hist(A, B, T), ref.mi_hist(A, B, T))
        assert oracle.mi_cost(A, B, T) == ref.mi_cost(A, B, T)


def test_sample_edges(oracle, ref):
    v = smooth_volume((6, 7, 8), 4, background=False)
    pts = [(0, 0, 0), (6.999, 5.5, 4.5), (7.0, 1, 1), (-0.0, 1, 1), (-1e-7, 1, 1), (3.5, 5.999, 4.999),
           (3.5, 6.0, 1), (float("nan"), 1, 1), (1e30, 1, 1), (2.25, 3.75, 1.5)]
    for p in pts:
        assert oracle.sample(v, *p) == ref.sample(v, *p), p


def test_quantise(oracle, ref):
    rng = np.random.default_rng(5)
    v = rng.normal(0, 300, (9, 10, 11)).astype(np.float32)
    assert np.array_equal(oracle.normalize(v), ref.normalize(v))
    assert np.array_equal(oracle.normalize(v, 255.99), ref.normalize(v, 255.99))


@pytest.mark.parametrize("shape", [(9, 10, 11), (16, 16, 16), (1, 3, 5), (2, 2, 2)])
def test_pyramid(oracle, ref, shape):
    v = np.random.default_rng(6).integers(0, 256, shape).astype(np.uint8)
    assert np.array_equal(oracle.downsample(v), ref.downsample(v))
    d = np.random.default_rng(7).standard_normal(shape + (3,)).astype(np.float32)
    out = tuple(2 * s for s in shape)
    assert np.array_equal(_bits(oracle.upsample(d, out)), _bits(ref.upsample(d, out)))


def _field(shape, seed, scale=2.0, zero_frac=0.2):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    d = gaussian_filter(r.standard_normal(shape + (3,)), (2, 2, 2, 0)).astype(np.float32) * scale
    d[r.random(shape) < zero_frac] = 0.0
    return d


@pytest.mark.parametrize("shape", [(12, 14, 16), (7, 11, 5)])
def test_cdm_steps(oracle, ref, shape):
    Is = smooth_volume(shape, 8, background=False)
    It = smooth_volume(shape, 9)
    d = _field(shape, 10, 3.0)
    Js = oracle.compose(Is, d)
    assert np.array_equal(Js, ref.compose(Is, d))
    for kind in "rd":
        a, ca = oracle.gradient(Js, It, kind)
        b, cb = ref.gradient(Js, It, kind)
        assert np.array_equal(_bits(a), _bits(b))
        assert abs(ca - cb) <= 1e-6 * max(abs(cb), 1e-30)
    nd = _field(shape, 11, 1.0, 0.0)
    p = oracle.poisson(nd)
    assert np.array_equal(_bits(p), _bits(ref.poisson(nd)))
    assert oracle.max_length(p) == ref.max_length(p)
    assert np.array_equal(_bits(oracle.accumulate(d, p)), _bits(ref.accumulate(d, p)))
    for s in (0.8, 0.1, 0.0):
        assert np.array_equal(_bits(oracle.smooth(p, d, s)), _bits(ref.smooth(p, d, s)))


def test_cdm_improved(oracle, ref):
    rng = np.random.default_rng(12)
    for n in range(0, 10):
        for _ in range(20):
            c = list(np.cumsum(rng.normal(-0.01, 0.02, n)).astype(np.float32))
            assert oracle.improved(c) == ref.improved(c)


def test_cdm_loop_body(oracle, ref):
    from test_gpu_cdm import oracle_level_iterations
    from scipy.ndimage import shift
    It = smooth_volume((26, 30, 22), 14, sigma=2.5)
    Is = shift(It.astype(np.float32), (0.5, 1.0, -1.0), order=1).astype(np.uint8)
    _, f_r = ref.cdm_iterations(It, Is, 5, want_field=True)
    assert np.array_equal(_bits(oracle_level_iterations(oracle, It, Is, 5)), _bits(f_r))


@pytest.mark.parametrize("shape,kind", [((36, 40, 32), "r"), ((33, 29, 35), "d"), ((20, 20, 20), "r")])
def test_whole_cdm(oracle, ref, shape, kind):
    from scipy.ndimage import shift
    It = smooth_volume(shape, 13, sigma=2.5)
    Is = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)
    ok_o, f_o, _ = oracle.cdm(It, Is, gradient_type=kind)
    ok_r, f_r = ref.cdm(It, Is, gradient_type=kind)
    assert ok_o == ok_r
    rms = float(np.sqrt(((f_o - f_r) ** 2).sum(-1).mean()))
    assert rms <= 1e-6, rms


@pytest.mark.parametrize("shape", [(13, 17, 11), (1, 1, 9), (9, 1, 1), (2, 3, 1), (20, 9, 14)])
@pytest.mark.parametrize("bandwidth", [0, 1, 2, 3])
def test_gaussian(oracle, ref, shape, bandwidth):
    rng = np.random.default_rng(bandwidth * 100 + sum(shape))
    img = (rng.standard_normal(shape) * 300).astype(np.float32)
    field = rng.standard_normal(shape + (3,)).astype(np.float32)
    assert np.array_equal(_bits(oracle.gaussian(img, bandwidth)), _bits(ref.gaussian(img, bandwidth)))
    assert np.array_equal(_bits(oracle.gaussian(field, bandwidth)),
                          _bits(ref.gaussian(field, bandwidth)))


@pytest.mark.parametrize("levels,iters,smoothing,bandwidth",
                         [(1, 5, 0.8, 0), (2, 4, 0.8, 0), (3, 3, 0.5, 2), (2, 3, 0.8, 8)])
def test_cdm_fixed_schedule(oracle, ref, levels, iters, smoothing, bandwidth):
    """The fixed-schedule CDM extension against the same schedule composed
    from the reference's own pyramid, loop-body and gaussian routines."""
    from scipy.ndimage import shift
    It = smooth_volume((26, 30, 22), 21, sigma=2.5)
    Is = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)
    fo, n = oracle.cdm_fixed(It, Is, levels, iters, 0.3, smoothing, bandwidth)
    fr = ref.cdm_fixed(It, Is, levels, iters, smoothing, bandwidth)
    assert np.array_equal(_bits(fo), _bits(fr))
    assert n == levels * iters


def _wfield(shape, seed, scale=3.0):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    return gaussian_filter(r.standard_normal(shape + (3,)), (1, 1, 1, 0)).astype(np.float32) * scale


WARP_SHAPES = [(9, 11, 13), (20, 17, 15), (3, 4, 5), (1, 6, 7)]


@pytest.mark.parametrize("shape", WARP_SHAPES)
def test_warping(oracle, ref, shape):
    """displacement_to_mapping (with and without T), mapping_to_displacement,
    inv_displacement_to_mapping, compose_mapping (uint8 and float) and both
    Jacobian determinants (numerical/dif.hpp:9-86, 356-465)."""
    rng = np.random.default_rng(sum(shape))
    d = _wfield(shape, 1)
    T = oracle.affine_matrix([1.5, -2, 0.7, 0.1, -0.05, 0.08, 1.05, 0.97, 1.02, 0.02, -0.01, 0.03],
                             np.array(shape[::-1], np.int32), [1, 1, 1],
                             np.array(shape[::-1], np.int32), [1.2, 1, 1])
    for TT in (None, T):
        assert np.array_equal(_bits(oracle.displacement_to_mapping(d, TT)),
                              _bits(ref.displacement_to_mapping(d, TT)))
    m = oracle.displacement_to_mapping(d, T)
    assert np.array_equal(_bits(oracle.mapping_to_displacement(m)), _bits(ref.mapping_to_displacement(m)))
    out_shape = (shape[0] + 2, max(1, shape[1] - 1), shape[2] + 3)
    assert np.array_equal(_bits(oracle.inv_displacement_to_mapping(d, T, out_shape)),
                          _bits(ref.inv_displacement_to_mapping(d, T, out_shape)))
    mm = m * np.float32(0.9)
    I8 = (rng.random(shape) * 255).astype(np.uint8)
    assert np.array_equal(oracle.compose_mapping(I8, mm), ref.compose_mapping(I8, mm))
    If = rng.standard_normal(shape).astype(np.float32)
    assert np.array_equal(_bits(oracle.compose_mapping(If, mm)), _bits(ref.compose_mapping(If, mm)))
    assert np.array_equal(_bits(oracle.jacobian_determinant_dis(d)), _bits(ref.jacobian_determinant_dis(d)))
    assert np.array_equal(_bits(oracle.jacobian_determinant(m)), _bits(ref.jacobian_determinant(m)))


def test_correlation_r(oracle, ref):
    """mm_reg's r: correlation_ygz and correlation (numerical/statistics.hpp:569-619)."""
    rng = np.random.default_rng(3)
    for shape in [(20, 17, 15), (1, 1, 7), (5, 5, 5)]:
        y = (rng.random(shape) * 255).astype(np.uint8)
        y[rng.random(shape) < 0.3] = 0
        x = (rng.random(shape) * 255).astype(np.uint8)
        for masked in (0, 1):
            assert oracle.correlation(x, y, masked) == ref.correlation(x, y, masked)
    z = np.zeros((4, 4, 4), np.uint8)
    assert oracle.correlation(z, z, 1) == ref.correlation(z, z, 1) == 0.0


@pytest.mark.parametrize("tshape,mshape", [((20, 22, 18), (24, 20, 26)), ((9, 10, 11), (5, 6, 7))])
def test_match_fov(oracle, ref, tshape, mshape):
    rng = np.random.default_rng(5)
    It = (rng.random(tshape) * 255).astype(np.uint8)
    arg = np.array([2, -3, 1.5, 0.1, 0.05, -0.08, 1.1, 0.95, 1.02, 0.01, 0, 0.02], np.float32)
    tvs, mvs = [1, 1, 1.2], [1.1, 1, 1]
    T = oracle.affine_matrix(arg, np.array(tshape[::-1], np.int32), tvs,
                             np.array(mshape[::-1], np.int32), mvs)
    assert np.array_equal(oracle.match_fov(It, T, mshape), ref.match_fov(It, tvs, mshape, mvs, arg))


@pytest.mark.parametrize("reverse,masked", [(True, True), (False, False)])
def test_mm_reg_pipeline(oracle, ref, reverse, masked):
    """The nonlinear half of the reference's mm_reg (skip_linear with a given
    affine: calculate_linear_r, It_match_fov, forward and reverse cdm,
    compute_mapping_from_displacement, calculate_nonlinear_r) against the same
    steps composed from the oracle (tests/mmreg_oracle.py)."""
    from scipy.ndimage import shift
    from mmreg_oracle import oracle_mm_reg
    It = smooth_volume((30, 34, 28), 3, sigma=2.5)
    I = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)[1:, :, 2:]
    arg = np.array([0.5, -0.3, 0.2, 0.02, -0.01, 0.015, 1.0, 1.0, 1.0, 0, 0, 0], np.float32)
    vs_t, vs_s = [1, 1, 1], [1.1, 1.0, 0.9]
    a = ref.mm_reg(It, vs_t, I, vs_s, arg, reverse=reverse, masked_r=masked)
    b = oracle_mm_reg(oracle, It, vs_t, I, vs_s, arg, reverse=reverse, masked_r=masked)
    keys = ["r", "J_lin", "It_fov", "t2f", "to2from", "JJ"] + (["f2t", "from2to"] if reverse else [])
    for k in keys:
        assert np.array_equal(_bits(a[k]), _bits(b[k])), k


def test_otsu_and_median(oracle, ref):
    """segmentation::otsu_threshold / otsu_median (numerical/otsu.hpp:66-108)
    and tipl::median's chunked rule (numerical/statistics.hpp:341-363)."""
    for img in prep_images():
        assert np.float32(oracle.otsu_threshold(img)) == np.float32(ref.otsu_threshold(img))
        assert np.float32(oracle.otsu_median(img)) == np.float32(ref.otsu_median(img))
        assert np.float32(oracle.median(img)) == np.float32(ref.median(img))


def test_median_full_size(oracle, ref):
    """The chunked median at 182x218x182 (7.2M values, 256 uneven chunks)."""
    rng = np.random.default_rng(77)
    v = (rng.gamma(2.0, 100.0, 182 * 218 * 182)).astype(np.float32)
    assert np.float32(oracle.median(v)) == np.float32(ref.median(v))


def test_subject_pre(oracle, ref):
    """reg::subject_image_pre (reg/mm_reg.hpp:14-20): gaussian, then
    normalize_otsu_median(255.99), then the uint8 conversion."""
    for img in prep_images():
        assert np.array_equal(oracle.subject_pre(img), ref.subject_pre(img)), img.shape


def test_quantise_int16(oracle, ref):
    """normalize_upper_lower2 on int16 input (numerical/numerical.hpp:909-927)."""
    rng = np.random.default_rng(8)
    for shape, lo, hi in [((9, 10, 11), -32768, 32767), ((5, 6, 7), 0, 4000), ((3, 4, 5), -7, 3)]:
        v = rng.integers(lo, hi, shape, endpoint=True).astype(np.int16)
        for upper in (255.99, 255.999):
            assert np.array_equal(oracle.normalize(v, upper), ref.normalize(v, upper))
    c = np.full((4, 4, 4), 12, np.int16)
    assert np.array_equal(oracle.normalize(c), ref.normalize(c))
