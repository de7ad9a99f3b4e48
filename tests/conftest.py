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
