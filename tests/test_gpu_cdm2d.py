"""GPU parity of the 2-D CDM path (cdm<out,uint8,2>, the reference's
cdm_cuda<out,uint8,2>, reg/cdm_cu.hpp:12-17), through the C ABI: every step
kernel bit-exact against the oracle on the same seeded inputs (the oracle's
2-D functions are pinned to the compiled reference in
test_oracle_cdm2d.py), and whole multi-level cdm() runs against the oracle's
field with the same iteration count (bar: <= 0.01 pixel RMS; in practice
bit-exact)."""
import numpy as np
import pytest

from test_oracle_cdm2d import field2, image2

pytestmark = pytest.mark.gpu

SHAPES = [(24, 28), (13, 7), (1, 33), (65, 130), (300, 257)]


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("shape", SHAPES)
def test_compose2d_bit_exact(ctx, oracle, shape):
    Is = image2(shape, 1, background=False)
    d = field2(shape, 2, 4.0)
    d[0, :3] = (np.nan, 0.5)  # a NaN coordinate samples as outside
    assert np.array_equal(ctx.compose2d(Is, d), oracle.compose2(Is, d))


@pytest.mark.parametrize("kind", ["r", "d"])
@pytest.mark.parametrize("shape", SHAPES)
def test_gradient2d_bit_exact(ctx, oracle, shape, kind):
    Js = image2(shape, 3)
    It = image2(shape, 4)
    nd, c = ctx.gradient2d(Js, It, kind)
    nd_o, c_o = oracle.gradient2(Js, It, kind)
    assert _bits_equal(nd, nd_o)
    assert c == c_o


@pytest.mark.parametrize("sweeps", [1, 2, 3, 5, 12])
@pytest.mark.parametrize("shape", SHAPES + [(6, 1100), (2, 1024), (700, 3), (9, 500), (11, 384)])
def test_poisson2d_bit_exact(ctx, oracle, shape, sweeps):
    """Four sweeps per launch up to 384 pixels wide, two up to 768, one
    beyond, over flat-index tiles: all bit-exact, odd counts included."""
    nd = field2(shape, 5, 1.0, 0.0)
    got, ml = ctx.poisson2d(nd, sweeps)
    want = nd.copy()
    oracle._c2_poisson(want, oracle._d2(want), sweeps)
    assert _bits_equal(got, want)
    assert ml == oracle.max_length2(want)


@pytest.mark.parametrize("shape", SHAPES)
def test_accumulate2d_bit_exact(ctx, oracle, shape):
    cur = field2(shape, 6, 2.0)
    nd = field2(shape, 7, 3.0)
    for k in (1.0, 0.37):
        want = oracle.accumulate2(cur, nd * np.float32(k))
        assert _bits_equal(ctx.accumulate2d(cur, nd, k), want)


@pytest.mark.parametrize("s", [0.8, 0.05, 0.0])
@pytest.mark.parametrize("shape", SHAPES)
def test_smooth2d_bit_exact(ctx, oracle, shape, s):
    d = field2(shape, 8, 2.0)
    dd = field2(shape, 9, 1.0)
    assert _bits_equal(ctx.smooth2d(d, dd, s), oracle.smooth2(d, dd, s))


@pytest.mark.parametrize("shape", SHAPES + [(2, 2), (3, 1)])
def test_pyramid2d_bit_exact(ctx, oracle, shape):
    v = np.random.default_rng(10).integers(0, 256, shape).astype(np.uint8)
    assert np.array_equal(ctx.downsample2d(v), oracle.downsample2(v))
    f = np.random.default_rng(11).standard_normal(shape + (2,)).astype(np.float32)
    out = tuple(2 * s for s in shape)
    assert _bits_equal(ctx.upsample2d(f, out), oracle.upsample2(f, out))


def _pair2(shape, seed, dx=1.5, dy=-2.0):
    from scipy.ndimage import shift
    It = image2(shape, seed, sigma=2.0)
    Is = shift(It.astype(np.float32), (dy, dx), order=1).astype(np.uint8)
    return It, Is


@pytest.mark.parametrize("shape", [(64, 72), (33, 70), (20, 20), (256, 256), (181, 217)])
def test_cdm2d_matches_oracle(ctx, oracle, shape):
    It, Is = _pair2(shape, 11)
    ok_o, f_o, it_o = oracle.cdm2(It, Is)
    rc, f, iters = ctx.cdm2d(It, Is)
    assert ok_o and rc == 0
    rms = float(np.sqrt(((f - f_o) ** 2).sum(-1).mean()))
    assert rms <= 0.01, rms
    assert iters == it_o


def test_cdm2d_d_gradient_and_two_modalities(ctx, oracle):
    shape = (96, 80)
    It, Is = _pair2(shape, 12)
    It2, Is2 = _pair2(shape, 13, -1.0, 1.0)
    for kind in ("d", "r"):
        ok_o, f_o, it_o = oracle.cdm2([It, It2], [Is, Is2], gradient_type=kind)
        p = ctx.cdm_param(gradient_type=kind)
        rc, f, iters = ctx.cdm2d([It, It2], [Is, Is2], p)
        assert ok_o and rc == 0
        assert float(np.sqrt(((f - f_o) ** 2).sum(-1).mean())) <= 0.01
        assert iters == it_o


def test_cdm2d_params(ctx, oracle):
    """resolution 1 (the top level runs), another speed, smoothing and
    min_dimension."""
    It, Is = _pair2((70, 90), 15)
    ok_o, f_o, it_o = oracle.cdm2(It, Is, resolution=1.0, speed=0.5, smoothing=0.2,
                                  min_dimension=24)
    p = ctx.cdm_param(resolution=1.0, speed=0.5, smoothing=0.2, min_dimension=24)
    rc, f, iters = ctx.cdm2d(It, Is, p)
    assert ok_o and rc == 0
    assert float(np.sqrt(((f - f_o) ** 2).sum(-1).mean())) <= 0.01
    assert iters == it_o


def test_cdm2d_zero_cost_returns_failed(ctx, oracle):
    import tipl_amd
    z = np.zeros((30, 30), np.uint8)
    ok_o, _, _ = oracle.cdm2(z, z)
    rc, _, _ = ctx.cdm2d(z, z)
    assert not ok_o and rc == tipl_amd.ERR_FAILED


def test_cdm2d_terminated_flag(ctx):
    import ctypes
    import tipl_amd
    It, Is = _pair2((40, 40), 14)
    flag = ctypes.c_uint8(1)
    rc, field, iters = ctx.cdm2d(It, Is, terminated=flag)
    assert rc == tipl_amd.ERR_TERMINATED and iters == 0 and not field.any()


def test_cdm2d_rejects_fixed_schedule(ctx):
    import tipl_amd
    It, Is = _pair2((32, 32), 16)
    p = ctx.cdm_param(levels=2, level_iterations=3)
    rc, _, _ = ctx.cdm2d(It, Is, p)
    assert rc == tipl_amd.ERR_INVALID
