"""Subject preprocessing on the GPU (tiplhip_volume_create_subject =
reg::subject_image_pre: binomial gaussian, normalize_otsu_median, uint8)
against the oracle, bit for bit: phantoms of several shapes and dtypes,
negative images, images with fewer than 2048 values above the Otsu
threshold (whole-set median) and the full 182x218x182 size (chunked
median)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,seed", [((20, 24, 28), 1), ((33, 17, 41), 2), ((64, 72, 60), 3),
                                        ((9, 10, 11), 4)])
def test_subject_pre_matches_oracle(ctx, oracle, shape, seed):
    from tipl_amd import phantoms
    f, m, _ = phantoms.serving_pair(seed, shape, False, np.float32)
    for img in (f, m, (50.0 - f).astype(np.float32), np.clip(np.rint(m), -32768, 32767).astype(np.int16)):
        print(f"subject_pre shape={shape} dtype={img.dtype}", flush=True)
        got = ctx.volume_subject(img).read()
        want = oracle.subject_pre(img.astype(np.float32))
        assert np.array_equal(got, want)


def test_subject_pre_sparse_foreground(ctx, oracle):
    """A few hundred bright voxels: fewer than 2048 values above the Otsu
    threshold, so the median is taken over all of them."""
    rng = np.random.default_rng(9)
    img = np.zeros((40, 40, 40), np.float32)
    idx = rng.integers(0, img.size, 300)
    img.reshape(-1)[idx] = rng.uniform(100, 900, 300).astype(np.float32)
    assert np.array_equal(ctx.volume_subject(img).read(), oracle.subject_pre(img))


def test_subject_pre_constant_image(ctx, oracle):
    img = np.full((16, 18, 20), 7.0, np.float32)
    assert np.array_equal(ctx.volume_subject(img).read(), oracle.subject_pre(img))


def test_subject_pre_full_size(ctx, oracle):
    from tipl_amd import phantoms
    fixed, moving, _ = phantoms.tissue_pair(1000)
    for img in (fixed, moving):
        assert np.array_equal(ctx.volume_subject(img).read(), oracle.subject_pre(img))
