"""tiplhip_cost_batch_noalloc (the cost evaluation that allocates no device
memory: its buffers live in a static device array) against
tiplhip_cost_batch, through the C ABI: costs bit-identical on two calls in a
row, the second on a smaller pair with fewer candidates (the array keeps the
first call's contents), for MI with 256 and 64 bins and for correlation
(including the role swap of a larger `from`); sizes beyond the array's
limits and an unpacked sampled volume are refused with an error."""
import numpy as np
import pytest

from conftest import random_params, smooth_volume

pytestmark = pytest.mark.gpu


def _T(p, fshape, tshape):
    import tipl_amd
    return tipl_amd.affine_matrix(p, fshape, (1, 1, 1), tshape, (1, 1, 1))


def _pair(ctx, fshape, tshape, seed, K):
    rng = np.random.default_rng(seed)
    A, B = smooth_volume(fshape, seed), smooth_volume(tshape, seed + 1)
    va, vb = ctx.volume(A), ctx.volume(B)
    Ts = np.stack([_T(p, fshape, tshape) for p in random_params(rng, K)])
    return va, vb, Ts


@pytest.mark.parametrize("cost,bins,K1", [("mi", 256, 40), ("mi", 64, 200), ("corr", 256, 256)])
def test_noalloc_equals_cost_batch(ctx, cost, bins, K1):
    big = _pair(ctx, (64, 72, 60), (60, 70, 66), 300, K1)
    small = _pair(ctx, (30, 34, 28), (26, 30, 32), 310, 8)
    for va, vb, Ts in (big, small):
        print(f"noalloc {cost} bins={bins} K={len(Ts)} from={va.shape}", flush=True)
        want = ctx.cost_batch(va, vb, Ts, cost, bins)        # also packs vb
        got = ctx.cost_batch_noalloc(va, vb, Ts, cost, bins)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # and again on the first pair after the smaller call
    va, vb, Ts = big
    assert np.array_equal(ctx.cost_batch_noalloc(va, vb, Ts, cost, bins).view(np.uint64),
                          ctx.cost_batch(va, vb, Ts, cost, bins).view(np.uint64))


def test_noalloc_correlation_role_swap(ctx):
    # from larger than to: correlation swaps them and samples `from`
    va, vb, Ts = _pair(ctx, (40, 44, 38), (20, 22, 24), 320, 12)
    want = ctx.cost_batch(va, vb, Ts, "corr")
    got = ctx.cost_batch_noalloc(va, vb, Ts, "corr")
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_noalloc_limits(ctx):
    import tipl_amd
    va, vb, Ts = _pair(ctx, (24, 26, 22), (24, 26, 22), 330, 4)
    fresh = ctx.volume(smooth_volume((24, 26, 22), 333))
    with pytest.raises(tipl_amd.TiplHipError, match="packed"):
        ctx.cost_batch_noalloc(va, fresh, Ts, "mi", 64)
    ctx.cost_batch(va, vb, Ts, "mi", 64)
    many = np.repeat(Ts[:1], 257, 0)
    with pytest.raises(tipl_amd.TiplHipError, match="256"):
        ctx.cost_batch_noalloc(va, vb, many, "mi", 64)
    with pytest.raises(tipl_amd.TiplHipError, match="scratch"):
        ctx.cost_batch_noalloc(va, vb, np.repeat(Ts[:1], 80, 0), "mi", 256)
    # at the 256-bin limit it still works
    got = ctx.cost_batch_noalloc(va, vb, np.repeat(Ts[:2], 40, 0)[:79], "mi", 256)
    want = ctx.cost_batch(va, vb, np.repeat(Ts[:2], 40, 0)[:79], "mi", 256)
    assert np.array_equal(got, want)
