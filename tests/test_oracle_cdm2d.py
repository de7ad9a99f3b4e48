"""Pins the oracle's 2-D CDM (cdm<out,uint8,2>, the reference's second GPU
instantiation, reg/cdm_cu.hpp:12-17) to the reference headers compiled into
oracle/_ref/libtiplref.so, bit for bit, and checks what the mathematics fixes
where a slip would not show against a single reference run: the bilinear
sampler's exact cases, the Poisson stencil's fixed point, a smoothing of a
constant field.  Golden vectors of the same functions are checked in
test_oracle_golden.py."""
import numpy as np
import pytest


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def image2(shape, seed, sigma=1.5, background=True):
    """A smooth random uint8 image with a zero margin."""
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    v = gaussian_filter(r.random(shape).astype(np.float32), sigma)
    v = (v - v.min()) / max(float(v.max() - v.min()), 1e-6) * 255.99
    v = v.astype(np.uint8)
    if background:
        m = np.zeros(shape, bool)
        m[tuple(slice(max(1, s // 8), s - max(1, s // 8)) for s in shape)] = True
        v[~m] = 0
    return v


def field2(shape, seed, scale=2.0, zero_frac=0.2):
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    d = gaussian_filter(r.standard_normal(shape + (2,)), (2, 2, 0)).astype(np.float32) * scale
    d[r.random(shape) < zero_frac] = 0.0
    return d


def test_sample2(oracle, ref):
    img = image2((9, 11), 1, background=False)
    pts = [(0.0, 0.0), (9.999, 7.999), (10.0, 3.0), (3.0, 8.0), (-0.0, 2.5), (-1e-7, 2.0),
           (4.25, 3.75), (float("nan"), 1.0), (1.0, float("inf")), (1e30, 1.0), (2.5, 0.5)]
    for x, y in pts:
        assert oracle.sample2(img, x, y) == ref.sample2(img, x, y), (x, y)
    # exact cases: a grid point returns its pixel; a midpoint of two pixels
    # whose sum is even returns their mean
    assert oracle.sample2(img, 3.0, 4.0) == img[4, 3]
    a, b = int(img[2, 5]), int(img[2, 6])
    if (a + b) % 2 == 0:
        assert oracle.sample2(img, 5.5, 2.0) == (a + b) // 2


@pytest.mark.parametrize("shape", [(9, 10), (16, 16), (1, 5), (3, 1), (2, 2), (17, 33)])
def test_pyramid2(oracle, ref, shape):
    v = np.random.default_rng(6).integers(0, 256, shape).astype(np.uint8)
    assert np.array_equal(oracle.downsample2(v), ref.downsample2(v))
    d = np.random.default_rng(7).standard_normal(shape + (2,)).astype(np.float32)
    out = tuple(2 * s for s in shape)
    assert np.array_equal(_bits(oracle.upsample2(d, out)), _bits(ref.upsample2(d, out)))
    # a constant image downsamples to itself
    c = np.full(shape, 77, np.uint8)
    assert (oracle.downsample2(c) == 77).all()


@pytest.mark.parametrize("shape", [(24, 28), (13, 7), (40, 33)])
def test_cdm2_steps(oracle, ref, shape):
    Is = image2(shape, 8, background=False)
    It = image2(shape, 9)
    d = field2(shape, 10, 3.0)
    Js = oracle.compose2(Is, d)
    assert np.array_equal(Js, ref.compose2(Is, d))
    for kind in "rd":
        a, ca = oracle.gradient2(Js, It, kind)
        b, cb = ref.gradient2(Js, It, kind)
        assert np.array_equal(_bits(a), _bits(b)), kind
        assert abs(ca - cb) <= 1e-6 * max(abs(cb), 1e-30)
    nd = field2(shape, 11, 1.0, 0.0)
    p = oracle.poisson2(nd)
    assert np.array_equal(_bits(p), _bits(ref.poisson2(nd)))
    assert oracle.max_length2(p) == ref.max_length2(p)
    assert np.array_equal(_bits(oracle.accumulate2(d, p)), _bits(ref.accumulate2(d, p)))
    for s in (0.8, 0.1, 0.0):
        assert np.array_equal(_bits(oracle.smooth2(p, d, s)), _bits(ref.smooth2(p, d, s)))


def test_cdm2_math(oracle):
    """Closed forms: compose with a zero field copies; compose with an
    integer shift moves pixels (0 where the cell leaves the image); a zero
    field accumulates to the current one; smoothing a constant field away
    from the array ends keeps it; Poisson sweeps of new_d = 0 stay 0; the SSD
    gradient of identical images is zero with zero cost; vector length is
    the 2-D norm."""
    shape = (12, 15)
    Is = image2(shape, 3, background=False)
    z = np.zeros(shape + (2,), np.float32)
    assert np.array_equal(oracle.compose2(Is, z), Is)
    sh = np.zeros_like(z)
    sh[..., 0] = 1.0  # x + 1
    J = oracle.compose2(Is, sh)
    # the last row and the last two columns have no bilinear cell (iy + 1 < h)
    assert np.array_equal(J[:-1, :-2], Is[:-1, 1:-1])
    assert not J[-1].any() and not J[:, -2:].any()
    cur = field2(shape, 4)
    assert np.array_equal(_bits(oracle.accumulate2(cur, z)), _bits(cur))
    c = np.full(shape + (2,), 0.5, np.float32)
    sm = oracle.smooth2(c, np.zeros_like(c), 0.8)
    assert np.allclose(sm[2:-2, 2:-2], 0.5, atol=1e-7)
    assert np.array_equal(oracle.poisson2(z), z)
    g, cost = oracle.gradient2(Is, Is, "d")
    assert cost == 0.0 and not g.any()
    f = np.zeros(shape + (2,), np.float32)
    f[3, 4] = (3.0, 4.0)
    assert oracle.max_length2(f) == 5.0


@pytest.mark.parametrize("shape,kind", [((64, 72), "r"), ((48, 40), "d"), ((20, 20), "r"),
                                        ((33, 70), "r")])
def test_whole_cdm2(oracle, ref, shape, kind):
    from scipy.ndimage import shift
    It = image2(shape, 13, sigma=2.0)
    Is = shift(It.astype(np.float32), (1.5, -2.0), order=1).astype(np.uint8)
    ok_o, f_o, n = oracle.cdm2(It, Is, gradient_type=kind)
    ok_r, f_r, _ = ref.cdm2(It, Is, gradient_type=kind)
    assert ok_o == ok_r
    assert n > 0
    rms = float(np.sqrt(((f_o - f_r) ** 2).sum(-1).mean()))
    assert rms <= 1e-6, rms


def test_whole_cdm2_two_modalities(oracle, ref):
    from scipy.ndimage import shift
    shape = (56, 60)
    It = [image2(shape, 21), image2(shape, 22)]
    Is = [shift(v.astype(np.float32), (-1.0, 1.5), order=1).astype(np.uint8) for v in It]
    ok_o, f_o, _ = oracle.cdm2(It, Is)
    ok_r, f_r, _ = ref.cdm2(It, Is)
    assert ok_o == ok_r
    rms = float(np.sqrt(((f_o - f_r) ** 2).sum(-1).mean()))
    assert rms <= 1e-6, rms


def test_whole_cdm2_zero_cost(oracle, ref):
    """Blank images give a zero cost: cdm() returns false."""
    z = np.zeros((30, 30), np.uint8)
    assert oracle.cdm2(z, z)[0] is False
    assert ref.cdm2(z, z)[0] is False
