"""GPU parity of post-registration warping (numerical/dif.hpp) and of the
device-resident mm_reg pipeline: every kernel bit-exact against the oracle,
and whole tiplhip_register runs (reverse CDM on the second stream, mappings,
r values, warped image) against the same steps composed from the oracle."""
import numpy as np
import pytest

from conftest import smooth_volume

pytestmark = pytest.mark.gpu

SHAPES = [(9, 11, 13), (20, 17, 15), (3, 4, 5), (1, 6, 7), (33, 70, 65)]


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _field(shape, seed, scale=3.0):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    return gaussian_filter(r.standard_normal(shape + (3,)), (1, 1, 1, 0)).astype(np.float32) * scale


def _T(oracle, shape, out_shape):
    return oracle.affine_matrix([1.5, -2, 0.7, 0.1, -0.05, 0.08, 1.05, 0.97, 1.02, 0.02, -0.01, 0.03],
                                np.array(shape[::-1], np.int32), [1, 1, 1],
                                np.array(out_shape[::-1], np.int32), [1.2, 1, 1])


@pytest.mark.parametrize("shape", SHAPES)
def test_mappings_bit_exact(ctx, oracle, shape):
    d = _field(shape, 1)
    T = _T(oracle, shape, shape)
    for TT in (None, T):
        assert np.array_equal(_bits(ctx.displacement_to_mapping(d, TT)),
                              _bits(oracle.displacement_to_mapping(d, TT)))
    m = oracle.displacement_to_mapping(d, T)
    assert np.array_equal(_bits(ctx.mapping_to_displacement(m)), _bits(oracle.mapping_to_displacement(m)))
    out_shape = (shape[0] + 2, max(1, shape[1] - 1), shape[2] + 3)
    assert np.array_equal(_bits(ctx.inv_displacement_to_mapping(d, T, out_shape)),
                          _bits(oracle.inv_displacement_to_mapping(d, T, out_shape)))


@pytest.mark.parametrize("shape", SHAPES)
def test_compose_mapping_bit_exact(ctx, oracle, shape):
    rng = np.random.default_rng(2)
    m = oracle.displacement_to_mapping(_field(shape, 3), _T(oracle, shape, shape)) * np.float32(0.9)
    src_shape = (shape[0] + 1, shape[1] + 2, max(1, shape[2] - 1))
    I8 = (rng.random(src_shape) * 255).astype(np.uint8)
    assert np.array_equal(ctx.compose_mapping(I8, m), oracle.compose_mapping(I8, m))
    If = rng.standard_normal(src_shape).astype(np.float32)
    assert np.array_equal(_bits(ctx.compose_mapping(If, m)), _bits(oracle.compose_mapping(If, m)))


@pytest.mark.parametrize("shape", SHAPES)
def test_jacobians_bit_exact(ctx, oracle, shape):
    d = _field(shape, 4, 1.0)
    assert np.array_equal(_bits(ctx.jacobian_determinant_dis(d)), _bits(oracle.jacobian_determinant_dis(d)))
    m = oracle.displacement_to_mapping(d)
    assert np.array_equal(_bits(ctx.jacobian_determinant(m)), _bits(oracle.jacobian_determinant(m)))


@pytest.mark.parametrize("shape", [(20, 22, 18), (7, 5, 3), (64, 64, 64), (1, 1, 17)])
def test_match_fov_and_correlation(ctx, oracle, shape):
    """It_match_fov in place on a device volume, then r (masked and not) over
    sizes with and without a 16-voxel tail."""
    rng = np.random.default_rng(6)
    It = (rng.random(shape) * 255).astype(np.uint8)
    It[rng.random(shape) < 0.2] = 0
    x = (rng.random(shape) * 255).astype(np.uint8)
    mshape = (shape[0] + 3, max(1, shape[1] - 2), shape[2] + 1)
    T = _T(oracle, shape, mshape)
    v = ctx.volume(It)
    ctx.match_fov(v, T, mshape)
    want = oracle.match_fov(It, T, mshape)
    assert np.array_equal(v.read(), want)
    s, s2, h = v.stats()
    assert s == int(want.astype(np.int64).sum()) and s2 == int((want.astype(np.int64) ** 2).sum())
    vx = ctx.volume(x)
    for masked in (True, False):
        assert ctx.correlation(vx, v, masked) == oracle.correlation(x, want, masked)


def _pipeline_case(seed=3):
    from scipy.ndimage import shift
    It = smooth_volume((30, 34, 28), seed, sigma=2.5)
    I = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)[1:, :, 2:]
    arg = np.array([0.5, -0.3, 0.2, 0.02, -0.01, 0.015, 1.0, 1.0, 1.0, 0, 0, 0], np.float32)
    return It, I, arg, (1.0, 1.0, 1.0), (1.1, 1.0, 0.9)


OUTPUTS = ("f2t", "to2from", "from2to", "warped")


@pytest.mark.parametrize("masked", [True, False])
def test_register_pipeline_fixed_schedule(ctx, oracle, masked):
    """tiplhip_register with a given affine, both CDM directions in the
    fixed-schedule extension (bit-exact end to end): every output equals the
    oracle's mm_reg composition bit for bit."""
    import tipl_amd
    from mmreg_oracle import oracle_mm_reg
    It, I, arg, vt, vs = _pipeline_case()
    res = tipl_amd.register(ctx, It, I, fixed_vs=vt, moving_vs=vs, nonlinear=True, reverse=True,
                            masked_r=masked, arg=arg, outputs=OUTPUTS, levels=2,
                            level_iterations=4, smoothing=0.5, field_bandwidth=2)
    want = oracle_mm_reg(oracle, It, vt, I, vs, arg, masked_r=masked, cdm_fixed=(2, 4, 0.5, 2))
    assert res["status"] == 0 and res["cdm_iterations"] == 8
    assert np.array_equal(_bits(res["T"]), _bits(want["T"]))
    assert res["r_linear"] == want["r"][0] and res["r_nonlinear"] == want["r"][1]
    for k, w in (("field", "t2f"), ("f2t", "f2t"), ("to2from", "to2from"), ("from2to", "from2to"),
                 ("warped", "JJ")):
        assert np.array_equal(_bits(res[k]), _bits(want[w])), k


def test_register_pipeline_reference_schedule(ctx, oracle):
    """The same with cdm()'s own schedule (cost-driven early exit): the fields
    agree with the oracle's to <= 0.01 voxel RMS (in practice exactly), the
    r values to 1e-4."""
    import tipl_amd
    from mmreg_oracle import oracle_mm_reg
    It, I, arg, vt, vs = _pipeline_case(4)
    res = tipl_amd.register(ctx, It, I, fixed_vs=vt, moving_vs=vs, nonlinear=True, reverse=True,
                            arg=arg, outputs=OUTPUTS)
    want = oracle_mm_reg(oracle, It, vt, I, vs, arg)
    assert res["status"] == 0
    assert res["r_linear"] == want["r"][0]
    for k, w in (("field", "t2f"), ("f2t", "f2t")):
        rms = float(np.sqrt(((res[k] - want[w]) ** 2).sum(-1).mean()))
        assert rms <= 0.01, (k, rms)
    assert abs(res["r_nonlinear"] - want["r"][1]) <= 1e-4
    assert res["r_nonlinear"] > res["r_linear"]


def test_register_affine_only_outputs(ctx, oracle):
    """An affine-only request reports calculate_linear_r's r and J."""
    import tipl_amd
    It, I, arg, vt, vs = _pipeline_case(5)
    res = tipl_amd.register(ctx, It, I, fixed_vs=vt, moving_vs=vs, arg=arg, outputs=("warped",))
    T = oracle.affine_matrix(arg, np.array(It.shape[::-1], np.int32), vt,
                             np.array(I.shape[::-1], np.int32), vs)
    J = oracle.resample(I, It.shape, T)
    assert np.array_equal(res["warped"], J)
    assert res["r_linear"] == np.float32(oracle.correlation(J, It, True))
    assert res["r_nonlinear"] == 0.0 and res["cdm_iterations"] == 0


def test_register_rejects_reverse_outputs_without_reverse(ctx):
    import tipl_amd
    It, I, arg, vt, vs = _pipeline_case(6)
    with pytest.raises(tipl_amd.TiplHipError):
        tipl_amd.register(ctx, It, I, nonlinear=True, arg=arg, outputs=("from2to",))


def test_register_files_equals_arrays(ctx, tmp_path):
    """Real-data entry: the same pair written as gzip NIfTI (one of them in
    RAS orientation with anisotropic voxels) registers exactly as the arrays
    the reader returns."""
    import tipl_amd
    from tipl_amd import phantoms
    f, m, _ = phantoms.serving_pair(4200, (40, 44, 36), True, np.float32)
    tipl_amd.nifti_write(tmp_path / "t.nii.gz", f, (1.0, 1.0, 1.0))
    tipl_amd.nifti_write(tmp_path / "s.nii.gz", m[:, ::-1, ::-1].copy(), (1.2, 1.0, 1.1),
                         srow=np.array([1.2, 0, 0, -10, 0, 1.0, 0, -20, 0, 0, 1.1, -5], np.float32))
    res = tipl_amd.register_files(ctx, tmp_path / "t.nii.gz", tmp_path / "s.nii.gz", nonlinear=True,
                                  reverse=True, outputs=("warped",))
    ff, fi = tipl_amd.nifti_read(tmp_path / "t.nii.gz")
    mm, mi = tipl_amd.nifti_read(tmp_path / "s.nii.gz")
    assert np.array_equal(mm, m) and mi["flip_swap"] == [0, 1]
    want = tipl_amd.register(ctx, ff, mm, fixed_vs=tuple(fi["vs"]), moving_vs=tuple(mi["vs"]),
                             preprocess="mm_reg", nonlinear=True, reverse=True, outputs=("warped",))
    assert res["status"] == 0 and np.array_equal(res["T"], want["T"])
    assert np.array_equal(res["field"], want["field"]) and np.array_equal(res["warped"], want["warped"])
    assert res["r_nonlinear"] == want["r_nonlinear"]
