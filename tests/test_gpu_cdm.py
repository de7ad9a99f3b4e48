"""GPU parity of the CDM path: every step kernel bit-exact against the oracle
on the same seeded inputs, and whole multi-level cdm() runs against the
oracle's field (bar: <= 0.01 voxel RMS; in practice bit-exact)."""
import numpy as np
import pytest

from conftest import smooth_volume

pytestmark = pytest.mark.gpu

SHAPES = [(20, 24, 28), (13, 17, 11), (1, 9, 33)]


def _field(shape, seed, scale=2.0, zero_frac=0.2):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    d = gaussian_filter(r.standard_normal(shape + (3,)), (2, 2, 2, 0)).astype(np.float32) * scale
    d[r.random(shape) < zero_frac] = 0.0
    return d


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("shape", SHAPES)
def test_compose_bit_exact(ctx, oracle, shape):
    Is = smooth_volume(shape, 1, background=False)
    d = _field(shape, 2, 4.0)
    assert np.array_equal(ctx.compose(Is, d), oracle.compose(Is, d))


@pytest.mark.parametrize("kind", ["r", "d"])
@pytest.mark.parametrize("shape", SHAPES)
def test_gradient_bit_exact(ctx, oracle, shape, kind):
    Js = smooth_volume(shape, 3)
    It = smooth_volume(shape, 4)
    nd, c = ctx.gradient(Js, It, kind)
    nd_o, c_o = oracle.gradient(Js, It, kind)
    assert _bits_equal(nd, nd_o)
    assert abs(c - c_o) <= 1e-6 * max(abs(c_o), 1e-30)


@pytest.mark.parametrize("shape", SHAPES)
def test_poisson_bit_exact(ctx, oracle, shape):
    """12 sweeps run as six two-sweep launches; 1, 2, 3 and 5 sweeps cover the
    first-sweep and odd-remainder paths."""
    nd = _field(shape, 5, 1.0, 0.0)
    for sweeps in (12, 1, 2, 3, 5):
        out, ml = ctx.poisson(nd, sweeps)
        want = oracle.poisson(nd, sweeps)
        assert _bits_equal(out, want), sweeps
        assert ml == oracle.max_length(want)


@pytest.mark.parametrize("shape", [(65, 9, 17), (1, 70, 3), (130, 2, 33), (3, 3, 40)])
def test_poisson_tile_edges(ctx, oracle, shape):
    """Shapes whose rows end inside a 64 x 8 tile, one-voxel-wide planes and
    deep thin volumes: the flat-index wrap across tile and plane edges."""
    nd = _field(shape, 9, 1.0, 0.0)
    out, ml = ctx.poisson(nd, 12)
    want = oracle.poisson(nd, 12)
    assert _bits_equal(out, want)
    assert ml == oracle.max_length(want)


@pytest.mark.parametrize("shape", SHAPES)
def test_accumulate_bit_exact(ctx, oracle, shape):
    cur = _field(shape, 6, 2.0, 0.1)
    nd = _field(shape, 7, 1.5, 0.3)
    k = np.float32(0.3) / np.float32(1.7)
    got = ctx.accumulate(cur, nd, float(k))
    want = oracle.accumulate(cur, (nd * k).astype(np.float32))
    assert _bits_equal(got, want)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("s", [0.8, 0.05, 0.0])
def test_smooth_bit_exact(ctx, oracle, shape, s):
    d = _field(shape, 8, 1.0, 0.0)
    dd = _field(shape, 9, 1.0, 0.0)
    assert _bits_equal(ctx.smooth(d, dd, s), oracle.smooth(d, dd, s))


@pytest.mark.parametrize("ishape,oshape", [((10, 12, 14), (20, 24, 28)), ((7, 9, 6), (13, 17, 11))])
def test_upsample_bit_exact(ctx, oracle, ishape, oshape):
    d = _field(ishape, 10, 1.0, 0.0)
    assert _bits_equal(ctx.upsample(d, oshape), oracle.upsample(d * np.float32(2.0), oshape))


def _pair(shape, seed):
    from scipy.ndimage import shift
    It = smooth_volume(shape, seed, sigma=2.5)
    Is = shift(It.astype(np.float32), (1.0, -1.5, 2.0), order=1).astype(np.uint8)
    return It, Is


@pytest.mark.parametrize("shape", [(40, 48, 40), (37, 45, 41), (24, 20, 18)])
def test_cdm_matches_oracle(ctx, oracle, shape):
    It, Is = _pair(shape, 11)
    ok_o, f_o, it_o = oracle.cdm(It, Is)
    rc, f, iters = ctx.cdm(ctx.volume(It), ctx.volume(Is))
    assert ok_o and rc == 0
    rms = float(np.sqrt(((f - f_o) ** 2).sum(-1).mean()))
    assert rms <= 0.01, rms
    assert iters == it_o


def test_cdm_d_gradient_and_two_modalities(ctx, oracle):
    shape = (36, 40, 32)
    It, Is = _pair(shape, 12)
    It2, Is2 = _pair(shape, 13)
    for kind in ("d", "r"):
        ok_o, f_o, _ = oracle.cdm([It, It2], [Is, Is2], gradient_type=kind)
        p = ctx.cdm_param(gradient_type=kind)
        rc, f, _ = ctx.cdm([ctx.volume(It), ctx.volume(It2)], [ctx.volume(Is), ctx.volume(Is2)], p)
        assert ok_o and rc == 0
        assert float(np.sqrt(((f - f_o) ** 2).sum(-1).mean())) <= 0.01


def test_cdm_zero_cost_returns_failed(ctx, oracle):
    import tipl_amd
    z = np.zeros((20, 20, 20), np.uint8)
    ok_o, _, _ = oracle.cdm(z, z)
    rc, _, _ = ctx.cdm(ctx.volume(z), ctx.volume(z))
    assert not ok_o and rc == tipl_amd.ERR_FAILED


def test_cdm_terminated_flag(ctx):
    import ctypes
    import tipl_amd
    It, Is = _pair((32, 32, 32), 14)
    flag = ctypes.c_uint8(1)
    rc, field, iters = ctx.cdm(ctx.volume(It), ctx.volume(Is), terminated=flag)
    assert rc == tipl_amd.ERR_TERMINATED and iters == 0 and not field.any()


def oracle_level_iterations(oracle, It, Is, iters, speed=0.3, smoothing=0.8):
    """The body of cdm()'s loop (reg/cdm.hpp:418-470) `iters` times from a
    zero field, no early exit, composed from the oracle's step functions."""
    cur = np.zeros(It.shape + (3,), np.float32)
    theta = 0.0
    for _ in range(iters):
        nd, _ = oracle.gradient(oracle.compose(Is, cur), It, "r")
        nd = oracle.poisson(nd)
        if theta == 0.0:
            theta = oracle.max_length(nd)
        if theta == 0.0:
            break
        k = np.float32(speed) / np.float32(theta)
        nd = oracle.accumulate(cur, (nd * k).astype(np.float32))
        cur = oracle.smooth(nd, cur, smoothing)
    return cur


def test_cdm_level_iterations_match_oracle(ctx, oracle):
    It, Is = _pair((30, 34, 26), 15)
    f = ctx.cdm_level_iterations(ctx.volume(It), ctx.volume(Is), 6, want_field=True)
    assert _bits_equal(f, oracle_level_iterations(oracle, It, Is, 6))


@pytest.mark.parametrize("shape,levels,iters,smoothing,bandwidth",
                         [((26, 30, 22), 3, 6, 0.8, 8), ((40, 44, 36), 2, 5, 0.8, 0),
                          ((33, 29, 41), 1, 4, 0.5, 2), ((48, 40, 44), 3, 4, 0.8, 8)])
def test_cdm_fixed_schedule_matches_oracle(ctx, oracle, shape, levels, iters, smoothing, bandwidth):
    """The fixed-schedule extension (levels x iterations, gaussian field
    smoothing): every step is bit-exact and nothing depends on the FP64 cost,
    so whole multi-level fields are bit-exact."""
    It, Is = _pair(shape, 16)
    p = ctx.cdm_param(levels=levels, level_iterations=iters, smoothing=smoothing,
                      field_bandwidth=bandwidth)
    rc, f, n = ctx.cdm(ctx.volume(It), ctx.volume(Is), p)
    fo, no = oracle.cdm_fixed(It, Is, levels, iters, 0.3, smoothing, bandwidth)
    assert rc == 0 and n == no == levels * iters
    assert _bits_equal(f, fo)


def test_cdm_fixed_schedule_two_modalities(ctx, oracle):
    It, Is = _pair((30, 26, 34), 17)
    It2, Is2 = _pair((30, 26, 34), 18)
    p = ctx.cdm_param(levels=2, level_iterations=4, field_bandwidth=8)
    rc, f, n = ctx.cdm([ctx.volume(It), ctx.volume(It2)], [ctx.volume(Is), ctx.volume(Is2)], p)
    fo, no = oracle.cdm_fixed([It, It2], [Is, Is2], 2, 4, 0.3, 0.05, 8)
    assert rc == 0 and n == no
    assert _bits_equal(f, fo)
