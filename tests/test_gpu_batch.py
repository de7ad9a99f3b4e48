"""Multi-pair launches (tiplhip_cost_batch_pairs): the candidates of many
image pairs of uneven sizes in one histogram / sums launch.  Every pair's
costs must equal, bit for bit, what the single-pair entry point returns for
that pair alone, and the oracle's costs (MI: the same joint histogram, so the
entropy within 1e-9 relative; correlation: exact sums)."""
import numpy as np
import pytest

from conftest import random_params, smooth_volume

pytestmark = pytest.mark.gpu


def _T(p, fshape, tshape):
    import tipl_amd
    return tipl_amd.affine_matrix(p, fshape, (1, 1, 1), tshape, (1, 1, 1))


def _pairs(rng, n, lo, hi, kmax):
    out = []
    for i in range(n):
        fs = tuple(int(v) for v in rng.integers(lo, hi, 3))
        ts = tuple(int(v) for v in rng.integers(lo, hi, 3))
        F = smooth_volume(fs, 100 + i)
        M = smooth_volume(ts, 200 + i)
        K = int(rng.integers(1, kmax + 1))
        P = random_params(rng, K)
        Ts = np.stack([_T(p, fs, ts) for p in P])
        out.append((F, M, Ts))
    return out


@pytest.mark.parametrize("cost,bins", [("mi", 256), ("mi", 64), ("corr", 256)])
def test_pairs_equal_single_and_oracle(ctx, oracle, cost, bins):
    rng = np.random.default_rng(11 if cost == "mi" else 12)
    data = _pairs(rng, 7, 9, 48, 12)
    vols = [(ctx.volume(F), ctx.volume(M), Ts) for F, M, Ts in data]
    print(f"pairs cost={cost} bins={bins} shapes={[(F.shape, M.shape, len(T)) for F, M, T in data]}",
          flush=True)
    got = ctx.cost_batch_pairs(vols, cost, bins)
    for (F, M, Ts), (vf, vm, _), g in zip(data, vols, got):
        single = ctx.cost_batch(vf, vm, Ts, cost, bins)
        assert np.array_equal(g.view(np.uint64), single.view(np.uint64))
        for k in range(len(Ts)):
            if cost == "mi":
                o = oracle.mi_cost(F, M, Ts[k], bins)
                assert abs(g[k] - o) <= 1e-9 * max(abs(o), 1e-12)
            else:
                o = oracle.corr_cost(F, M, Ts[k])
                assert abs(g[k] - o) <= 1e-9 * abs(o) + 1e-12


def test_pairs_repeat_and_identity_crop(ctx):
    """A pair given twice, and integer-shift candidates (the correlation crop
    path) mixed with general ones, in one launch."""
    F = smooth_volume((30, 34, 28), 1)
    M = smooth_volume((30, 34, 28), 2)
    vf, vm = ctx.volume(F), ctx.volume(M)
    rng = np.random.default_rng(3)
    P = random_params(rng, 6)
    P[0] = [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0]
    P[1] = [2, -1, 3, 0, 0, 0, 1, 1, 1, 0, 0, 0]
    Ts = np.stack([_T(p, F.shape, M.shape) for p in P])
    for cost in ("corr", "mi"):
        a, b = ctx.cost_batch_pairs([(vf, vm, Ts), (vf, vm, Ts[::-1].copy())], cost)
        assert np.array_equal(a, b[::-1])
        assert np.array_equal(a, ctx.cost_batch(vf, vm, Ts, cost))


def test_pairs_volumes_from_another_context(ctx):
    import tipl_amd
    F = smooth_volume((20, 22, 24), 5)
    M = smooth_volume((24, 22, 20), 6)
    Ts = np.stack([_T(p, F.shape, M.shape) for p in random_params(np.random.default_rng(4), 5)])
    with tipl_amd.Context(0) as other:
        vf, vm = other.volume(F), other.volume(M)
        expect = other.cost_batch(vf, vm, Ts, "mi")
        (got,) = ctx.cost_batch_pairs([(vf, vm, Ts)], "mi")
        assert np.array_equal(got, expect)
        vf.free()
        vm.free()


def test_pairs_uneven_cubes(ctx):
    """Config-3-like: cubes from 64^3 to 128^3, several candidates each, one launch."""
    rng = np.random.default_rng(21)
    vols = []
    for i in range(8):
        s = int(rng.integers(64, 129))
        F = smooth_volume((s, s, s), 300 + i, sigma=3.0)
        M = smooth_volume((s, s, s), 400 + i, sigma=3.0)
        Ts = np.stack([_T(p, F.shape, M.shape) for p in random_params(rng, 8)])
        vf, vm = ctx.volume(F), ctx.volume(M)
        vols.append((vf, vm, Ts))
    for cost, bins in (("mi", 64), ("mi", 256), ("corr", 256)):
        got = ctx.cost_batch_pairs(vols, cost, bins)
        for (vf, vm, Ts), g in zip(vols, got):
            assert np.array_equal(g, ctx.cost_batch(vf, vm, Ts, cost, bins))


def test_pairs_rejects_bad_arguments(ctx):
    import tipl_amd
    F = smooth_volume((10, 10, 10), 7)
    vf = ctx.volume(F)
    T = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32), (2100, 1))
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.cost_batch_pairs([(vf, vf, T), (vf, vf, T)], "mi")
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.cost_batch_pairs([], "mi")
    with pytest.raises(tipl_amd.TiplHipError):
        ctx.cost_batch_pairs([(vf, vf, T[:3])], "mi", 128)


@pytest.mark.parametrize("cost,bins", [("mi", 64), ("mi", 256), ("corr", 256)])
def test_autotune_keeps_results(ctx, cost, bins):
    """The tuned launch shape is used for later launches of the same key and
    changes no result."""
    rng = np.random.default_rng(31)
    F = smooth_volume((40, 44, 36), 41)
    M = smooth_volume((44, 40, 38), 42)
    vf, vm = ctx.volume(F), ctx.volume(M)
    Ts = np.stack([_T(p, F.shape, M.shape) for p in random_params(rng, 12)])
    before = ctx.cost_batch(vf, vm, Ts, cost, bins)
    zc, pc, ms = ctx.autotune(vf, vm, Ts, cost, bins, reps=2)
    print(f"autotune {cost}{bins}: planes {zc}, workgroups/CU {pc}, {ms:.4f} ms")
    assert zc in (8, 16, 32) and pc in (0, 1, 2, 4, 8, 16, 32, 48, 64) and ms > 0
    after = ctx.cost_batch(vf, vm, Ts, cost, bins)
    assert np.array_equal(before, after)


def test_kernel_timing_counts_every_launch(ctx):
    """Kernel timing holds event pairs without waiting on them and resolves
    them when read and every 512 launches: 600 timed launches all count,
    results are unchanged, and re-enabling starts from zero."""
    rng = np.random.default_rng(32)
    F = smooth_volume((24, 26, 20), 43)
    vf = ctx.volume(F)
    Ts = np.stack([_T(p, F.shape, F.shape) for p in random_params(rng, 4)])
    want = ctx.cost_batch(vf, vf, Ts, "mi", 64)
    ctx.set_kernel_timing(True)
    for _ in range(600):
        ctx.cost_batch(vf, vf, Ts, "mi", 64)
    ms, n = ctx.kernel_time()
    assert n == 600 and ms > 0
    got = ctx.cost_batch(vf, vf, Ts, "mi", 64)
    assert ctx.kernel_time()[1] == 601
    ctx.set_kernel_timing(False)
    assert np.array_equal(got, want)
    ctx.set_kernel_timing(True)
    assert ctx.kernel_time() == (0.0, 0)
    ctx.set_kernel_timing(False)
