import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bindings import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.bindings import load_reference
    r = load_reference()
    if r is None:
        pytest.skip("compiled reference (oracle/_ref) not available")
    return r


@pytest.fixture(scope="session")
def ctx():
    import tipl_amd
    c = tipl_amd.Context(0)
    yield c
    c.close()


def smooth_volume(shape, seed, sigma=2.0, background=True):
    """A smooth random uint8 volume with a zero background margin."""
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed)
    v = gaussian_filter(r.random(shape).astype(np.float32), sigma)
    v = (v - v.min()) / max(float(v.max() - v.min()), 1e-6) * 255.99
    v = v.astype(np.uint8)
    if background:
        m = np.zeros(shape, bool)
        sl = tuple(slice(max(1, s // 8), s - max(1, s // 8)) for s in shape)
        m[sl] = True
        v[~m] = 0
    return v


def random_params(rng, K, t=3.0, r=0.2, s=0.1, a=0.05):
    P = np.zeros((K, 12), np.float32)
    P[:, 0:3] = rng.uniform(-t, t, (K, 3))
    P[:, 3:6] = rng.uniform(-r, r, (K, 3))
    P[:, 6:9] = rng.uniform(1 - s, 1 + s, (K, 3))
    P[:, 9:12] = rng.uniform(-a, a, (K, 3))
    return P


def prep_images():
    """Seeded float32 images for the preprocessing pins: tissue-like images
    with a zero background (fewer than half the voxels positive), images with
    a positive floor (more than half positive, so normalize_otsu_median
    subtracts upper*0.1), negative values, a constant image, a bimodal
    histogram and sizes on both sides of tipl::median's 2048-value chunking."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(2024)
    out = []
    for shape in [(20, 22, 18), (7, 9, 11), (33, 12, 5), (2, 32, 32), (1, 1, 2047), (1, 1, 2049)]:
        v = gaussian_filter(rng.random(shape), 1.5).astype(np.float32) * 900
        out.append(np.where(rng.random(shape) < 0.55, 0, v).astype(np.float32))
        out.append((v + 50).astype(np.float32))
    out.append((rng.standard_normal((10, 11, 12)) * 300).astype(np.float32))
    out.append(np.full((6, 6, 6), 42.0, np.float32))
    bi = np.where(rng.random((16, 16, 16)) < 0.4, 100.0, 700.0) + rng.normal(0, 5, (16, 16, 16))
    out.append(bi.astype(np.float32))
    # exactly half the voxels positive after the [1 2 1] pass: the boundary
    # of normalize_otsu_median's "more than half" rule
    half = np.zeros((1, 1, 16), np.float32)
    half[..., 1:7] = [100, 300, 500, 700, 400, 200]
    out.append(half)
    return out
