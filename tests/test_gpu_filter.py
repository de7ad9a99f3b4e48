"""filter::gaussian on the GPU (tiplhip_gaussian) against the oracle, bit for
bit: scalar images and vector<3> fields, ragged and degenerate shapes, the
bandwidths the reference's callers use (1 in subject preprocessing; 8, i.e.
sigma 2, for field smoothing), and the full 182x218x182 shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("shape", [(13, 17, 11), (1, 1, 9), (9, 1, 1), (1, 7, 1), (2, 3, 1),
                                   (33, 70, 130), (17, 9, 65)])
@pytest.mark.parametrize("bandwidth", [0, 1, 2, 3])
def test_gaussian_matches_oracle(ctx, oracle, shape, bandwidth):
    print(f"gaussian shape={shape} bandwidth={bandwidth}", flush=True)
    rng = np.random.default_rng(bandwidth * 1000 + sum(shape))
    img = (rng.standard_normal(shape) * 300).astype(np.float32)
    field = rng.standard_normal(shape + (3,)).astype(np.float32)
    assert np.array_equal(_bits(ctx.gaussian(img, bandwidth)), _bits(oracle.gaussian(img, bandwidth)))
    assert np.array_equal(_bits(ctx.gaussian(field, bandwidth)),
                          _bits(oracle.gaussian(field, bandwidth)))


def test_gaussian_field_sigma2(ctx, oracle):
    rng = np.random.default_rng(5)
    field = (rng.standard_normal((40, 48, 40, 3)) * 4).astype(np.float32)
    assert np.array_equal(_bits(ctx.gaussian(field, 8)), _bits(oracle.gaussian(field, 8)))


def test_gaussian_full_size(ctx, oracle):
    from tipl_amd import phantoms
    fixed, _, _ = phantoms.tissue_pair(1000)
    img = np.ascontiguousarray(fixed, np.float32)
    assert img.shape == (182, 218, 182)
    assert np.array_equal(_bits(ctx.gaussian(img, 1)), _bits(oracle.gaussian(img, 1)))


def test_gaussian_rejects_bad_arguments(ctx):
    import tipl_amd
    img = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.gaussian(img, 17)
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.gaussian(np.zeros((4, 4, 4, 2), np.float32), 1)
