"""Parity at the workloads' own sizes.

* 182x218x182 tissue pair, 64 candidates, 64x64 bins at the autotuned launch
  shape (the headline launch): every histogram counts each voxel once, its
  rows are the 64-bin value histogram of `from`, and each 64-bin bin is the
  sum of its 4x4 block of the 256-bin histogram of the same candidate (the
  256-bin kernel is the reference's histogram, pinned to the oracle); three
  candidates are bit-exact against the oracle; every candidate's cost agrees
  with the entropy of its own histogram computed here in FP64.
* 160x192x160 RBF-warped pair under the fixed schedule (3 levels, field
  smoothed by a sigma-2 binomial gaussian): the first two iterations of every
  level bit-exact against the oracle; the full 3 x 48 run gives a finite field
  and an image correlation that rises from 12 iterations per level on and ends
  above its starting value.
* Served requests at 182x218x182 and 256^3 (affine+CDM) equal the plain
  single-pair path bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COST_RTOL = 1e-9


@pytest.fixture(scope="module")
def tissue_dev(ctx):
    """The headline pair quantised on the device, as the benchmark uploads it."""
    import tipl_amd
    from tipl_amd import phantoms
    fixed, moving, _ = phantoms.tissue_pair(1000)
    vf, vm = ctx.volume(fixed, 255.99), ctx.volume(moving, 255.99)
    rng = np.random.default_rng(1000)
    P = phantoms.candidate_params(rng, 64)
    Ts = np.stack([tipl_amd.affine_matrix(p, fixed.shape, (1, 1, 1), moving.shape, (1, 1, 1))
                   for p in P])
    yield vf, vm, vf.read(), vm.read(), Ts
    vf.free()
    vm.free()


def _entropy_cost(H, bins):
    """get_mutual_info_mean's cost from a joint histogram, in FP64."""
    H = H.astype(np.float64)
    fh, th = H.sum(1), H.sum(0)
    nz = H > 0
    r, c = np.nonzero(nz)
    mu = H[nz]
    return -float((mu * np.log(mu / fh[r] / th[c])).sum() / float(bins * bins))


def test_headline_launch_64_bins_full_size(ctx, oracle, tissue_dev):
    vf, vm, F, M, Ts = tissue_dev
    zc, pc, _ = ctx.autotune(vf, vm, Ts, "mi", 64, reps=2)
    print(f"autotuned shape: {zc} planes per item, {pc} workgroups per CU", flush=True)
    H64 = ctx.mi_hist(vf, vm, Ts, bins=64)          # the tuned shape (same key)
    H256 = ctx.mi_hist(vf, vm, Ts, bins=256)
    fh64 = np.bincount((F >> 2).ravel(), minlength=64)
    for k in range(len(Ts)):
        assert H64[k].sum() == F.size
        assert np.array_equal(H64[k].sum(1), fh64), k
        blocks = H256[k].reshape(64, 4, 64, 4).sum(axis=(1, 3))
        assert np.array_equal(H64[k], blocks), k
    for k in (0, 31, 63):
        assert np.array_equal(H64[k], oracle.mi_hist(F, M, Ts[k], bins=64)), k
    costs = ctx.cost_batch(vf, vm, Ts, "mi", 64)
    for k in range(len(Ts)):
        want = _entropy_cost(H64[k], 64)
        assert abs(costs[k] - want) <= COST_RTOL * abs(want), k
    for k in (0, 63):
        want = oracle.mi_cost(F, M, Ts[k], bins=64)
        assert abs(costs[k] - want) <= COST_RTOL * abs(want), k


def test_headline_correlation_full_size(ctx, oracle, tissue_dev):
    vf, vm, F, M, Ts = tissue_dev
    costs = ctx.cost_batch(vf, vm, Ts, "corr")
    for k in (0, 40):
        assert costs[k] == oracle.corr_cost(F, M, Ts[k]), k


@pytest.fixture(scope="module")
def rbf_pair(ctx):
    from tipl_amd import phantoms
    fixed, moving, _ = phantoms.warped_pair(2000, (160, 192, 160))
    vt, vs = ctx.volume(fixed, 255.99), ctx.volume(moving, 255.99)
    yield vt, vs, vt.read(), vs.read()
    vt.free()
    vs.free()


def test_cdm_fixed_schedule_first_iterations_full_size(ctx, oracle, rbf_pair):
    vt, vs, It, Is = rbf_pair
    p = ctx.cdm_param(levels=3, level_iterations=2, field_bandwidth=8)
    rc, f, n = ctx.cdm(vt, vs, p)
    fo, no = oracle.cdm_fixed(It, Is, 3, 2, 0.3, 0.05, 8)
    assert rc == 0 and n == no == 6
    assert np.array_equal(f.view(np.uint32), fo.view(np.uint32))


def test_cdm_fixed_schedule_full_run(ctx, rbf_pair):
    vt, vs, It, Is = rbf_pair
    It64 = It.astype(np.float64).ravel()

    def r_of(J):
        return float(np.corrcoef(It64, J.astype(np.float64).ravel())[0, 1])

    rs = [r_of(Is)]
    for iters in (6, 12, 24, 48):
        p = ctx.cdm_param(levels=3, level_iterations=iters, field_bandwidth=8)
        rc, f, n = ctx.cdm(vt, vs, p)
        assert rc == 0 and n == 3 * iters
        assert np.isfinite(f).all()
        rs.append(r_of(ctx.compose(Is, f)))
    print("image r by iterations per level (0, 6, 12, 24, 48):", rs)
    # the first few iterations of a level can overshoot (measured: r 0.99637
    # before, 0.99631 after 6 per level); from 12 per level on r rises, and
    # the full run ends above where it started
    assert all(b >= a for a, b in zip(rs[2:], rs[3:])), rs
    assert rs[-1] > rs[0], rs


def _same(a, b):
    assert a["status"] == b["status"] == 0
    assert np.array_equal(a["T"].view(np.uint32), b["T"].view(np.uint32))
    assert a["cost"] == b["cost"] and a["evaluations"] == b["evaluations"]
    assert a["cdm_iterations"] == b["cdm_iterations"]
    assert a["r_linear"] == b["r_linear"] and a["r_nonlinear"] == b["r_nonlinear"]
    if a["field"] is not None:
        assert np.array_equal(a["field"].view(np.uint32), b["field"].view(np.uint32))


def test_served_equals_plain_full_sizes(ctx):
    import tipl_amd
    from tipl_amd import phantoms
    reqs = [phantoms.serving_pair(4100, (182, 218, 182), True, np.float32)[:2],
            phantoms.serving_pair(4101, (256, 256, 256), True, np.int16)[:2],
            phantoms.serving_pair(4102, (96, 96, 96), False, np.float32)[:2]]
    kinds = [True, True, False]
    plain = []
    for (f, m), nl in zip(reqs, kinds):
        print(f"plain {f.shape} nonlinear={nl}", flush=True)
        plain.append(tipl_amd.register(ctx, f, m, nonlinear=nl))
    with tipl_amd.Server(0, slots=4, hbm_budget=64 << 30) as srv:
        tickets = [srv.submit(f, m, nonlinear=nl) for (f, m), nl in zip(reqs, kinds)]
        served = [srv.wait(t) for t in tickets]
    for a, b in zip(served, plain):
        _same(a, b)
