"""HIP graph replay of cost launches (tiplhip_ctx_set_graphs).

A launch shape seen before on a context replays the graph captured the first
time: the upload of transforms and segment table, the histogram or sums
kernel (and the z-product table of a large launch), the entropy or finishing
kernels.  Every replay must give the eager path's costs bit for bit, with
eager work, other shapes and synchronisations between replays, and after the
buffers the graphs were captured on have been reallocated."""
import numpy as np
import pytest

from conftest import random_params, smooth_volume

pytestmark = pytest.mark.gpu


def _T(oracle, p, fshape, tshape):
    return oracle.affine_matrix(p, np.array(fshape[::-1], np.int32), (1, 1, 1),
                                np.array(tshape[::-1], np.int32), (1, 1, 1))


def _bits(a):
    return np.asarray(a).view(np.uint64)


@pytest.fixture
def gctx():
    import tipl_amd
    c = tipl_amd.Context(0)
    yield c
    c.close()


def test_replays_match_eager(gctx, oracle):
    rng = np.random.default_rng(41)
    A, B = smooth_volume((30, 34, 26), 42), smooth_volume((28, 30, 32), 43)
    va, vb = gctx.volume(A), gctx.volume(B)
    cases = [("mi", 256, 12), ("mi", 64, 12), ("corr", 256, 12), ("mi", 256, 5)]
    rounds = 4
    Ts = {c: [np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, c[2])])
              for _ in range(rounds)] for c in cases}
    eager = {c: [gctx.cost_batch(va, vb, T, cost=c[0], bins=c[1]) for T in Ts[c]] for c in cases}
    gctx.set_graphs(True)
    for r in range(rounds):
        for c in cases:
            got = gctx.cost_batch(va, vb, Ts[c][r], cost=c[0], bins=c[1])  # r = 0 captures
            assert np.array_equal(_bits(got), _bits(eager[c][r])), (c, r)
        # eager work and a synchronisation between replays
        gctx.set_kernel_timing(True)
        again = gctx.cost_batch(va, vb, Ts[cases[0]][r], cost="mi", bins=256)
        gctx.set_kernel_timing(False)
        assert np.array_equal(_bits(again), _bits(eager[cases[0]][r]))
        gctx.mi_hist(va, vb, Ts[cases[1]][r], bins=256)  # leaves the histogram buffer dirty
        gctx.synchronize()
    want = np.array([oracle.mi_cost(A, B, T, bins=256) for T in Ts[cases[0]][rounds - 1]])
    np.testing.assert_allclose(eager[cases[0]][rounds - 1], want, rtol=1e-9)


def test_replays_after_reallocation_and_multi_pair(gctx, oracle):
    rng = np.random.default_rng(44)
    shapes = [((20, 22, 24), (22, 20, 24)), ((16, 18, 20), (18, 16, 22)), ((26, 20, 18), (24, 22, 20))]
    vols = [(smooth_volume(f, 45 + i), smooth_volume(t, 55 + i)) for i, (f, t) in enumerate(shapes)]
    dv = [(gctx.volume(f), gctx.volume(t)) for f, t in vols]

    def pairs(seed):
        r = np.random.default_rng(seed)
        return [(vf, vt, np.stack([_T(oracle, p, f.shape, t.shape) for p in random_params(r, 7)]))
                for (vf, vt), (f, t) in zip(dv, vols)]

    sets = [pairs(100 + i) for i in range(3)]
    eager = [gctx.cost_batch_pairs(s, "mi", 256) for s in sets]
    gctx.set_graphs(True)
    for i, s in enumerate(sets):
        got = gctx.cost_batch_pairs(s, "mi", 256)
        for a, b in zip(got, eager[i]):
            assert np.array_equal(_bits(a), _bits(b))
    # a much larger launch reallocates the buffers the graphs were captured on
    big = np.stack([_T(oracle, p, vols[0][0].shape, vols[0][1].shape) for p in random_params(rng, 300)])
    gctx.cost_batch(dv[0][0], dv[0][1], big, cost="mi", bins=256)
    for i, s in enumerate(sets):
        got = gctx.cost_batch_pairs(s, "mi", 256)
        for a, b in zip(got, eager[i]):
            assert np.array_equal(_bits(a), _bits(b))
    gctx.set_graphs(False)
    for i, s in enumerate(sets):
        got = gctx.cost_batch_pairs(s, "mi", 256)
        for a, b in zip(got, eager[i]):
            assert np.array_equal(_bits(a), _bits(b))


def test_large_launch_with_z_table_replays(gctx, oracle):
    """A launch of at least 2^25 samples captures the z-product table kernel too."""
    rng = np.random.default_rng(47)
    A, B = smooth_volume((64, 64, 64), 48), smooth_volume((64, 64, 64), 49)
    va, vb = gctx.volume(A), gctx.volume(B)
    sets = [np.stack([_T(oracle, p, A.shape, B.shape) for p in random_params(rng, 130)]) for _ in range(3)]
    eager = [gctx.cost_batch(va, vb, T, cost=c, bins=64) for T in sets for c in ("mi", "corr")]
    gctx.set_graphs(True)
    for rep in range(2):
        got = [gctx.cost_batch(va, vb, T, cost=c, bins=64) for T in sets for c in ("mi", "corr")]
        for a, b in zip(got, eager):
            assert np.array_equal(_bits(a), _bits(b)), rep
