"""Pins the host logic of the C++ mirror (include/tipl/optimization.hpp,
include/tipl/reg/linear.hpp) against the reference's own routines, compiled
from its headers (oracle/_ref), on the CPU:

  * line_search and lbfgs (numerical/optimization.hpp:137-512) on quadratic
    bowls: the same candidates in the same order, the same result bit for
    bit, and the known minimum (inside the bounds) or the bound (outside);
  * has_mask and estimate_affine_param (numerical/basic_op.hpp:811-830,
    numerical/transformation.hpp:933-1021), bit for bit;
  * the whole linear() search (reg/linear.hpp:290-700: bounds, stage list,
    pyramid, modality mean / 2, line search then L-BFGS, the mask-based
    start and its voxel-size adjustment) on config-0 blob pairs with the
    oracle's costs: the same evaluation count, parameters and cost bit for
    bit as the reference's runner driven in the same order.

The reference races line search against L-BFGS on two threads
(reg/linear.hpp:446-499); its driver here runs them one after the other,
which is the mirror's schedule, so the comparison is deterministic."""
import numpy as np
import pytest

from host_shims import Mirror, OracleCosts, Recorder, ReferenceSchedule

# reg_bound-like ranges for the 12 affine parameters (t, r, s, a)
UPPER = np.array([30, 30, 30, 0.6, 0.4, 0.4, 1.5, 1.5, 1.5, 0.1, 0.1, 0.1], np.float32)
LOWER = np.array([-30, -30, -30, -0.6, -0.4, -0.4, 0.7, 0.7, 0.7, -0.1, -0.1, -0.1], np.float32)
STRATEGY = [0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1]


@pytest.fixture(scope="module")
def mirror():
    return Mirror()


@pytest.fixture(scope="module")
def rsched(ref):
    return ReferenceSchedule(ref)


def bowl(center, weights, cross=0.0):
    c = np.asarray(center, np.float64)
    w = np.asarray(weights, np.float64)

    def f(x):
        d = x.astype(np.float64) - c
        return float((w * d * d).sum() + cross * d[0] * d[3])
    return f


BOWLS = [
    # centre inside the bounds, different curvature per parameter
    (np.array([4.0, -7.5, 2.25, 0.1, -0.05, 0.2, 1.1, 0.9, 1.05, 0.02, -0.03, 0.01]),
     np.array([1, 2, 0.5, 40, 80, 60, 30, 50, 20, 300, 200, 500]), 0.0),
    # centre outside the bounds in three parameters, coupled terms
    (np.array([45.0, -3.0, 1.0, 0.8, 0.0, -0.1, 1.2, 2.0, 1.0, 0.0, 0.5, -0.02]),
     np.array([0.3, 1, 1, 10, 10, 10, 10, 10, 10, 100, 100, 100]), 0.7),
]


@pytest.mark.parametrize("case", range(len(BOWLS)))
@pytest.mark.parametrize("grid", [16, 5])
def test_line_search_bowl(mirror, rsched, case, grid):
    center, w, cross = BOWLS[case]
    f = bowl(center, w, cross)
    x0 = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
    up, lo = UPPER.copy(), LOWER.copy()
    up[11] = lo[11] = 0.0          # one fixed parameter, as the stage bounds make them
    ra, rb = Recorder(f), Recorder(f)
    xa, va = mirror.line_search(ra, x0, up, lo, STRATEGY, grid, f(x0))
    xb, vb = rsched.line_search(rb, x0, up, lo, STRATEGY, grid, f(x0))
    assert np.array_equal(ra.candidates(), rb.candidates())
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)) and va == vb
    assert va < f(x0)
    assert np.all(xa <= up) and np.all(xa >= lo)


@pytest.mark.parametrize("case", range(len(BOWLS)))
def test_lbfgs_bowl(mirror, rsched, case):
    center, w, cross = BOWLS[case]
    f = bowl(center, w, cross)
    x0 = np.array([1.5, 2, -1, 0.05, 0.05, 0, 1, 1, 1, 0, 0, 0], np.float32)
    ra, rb = Recorder(f), Recorder(f)
    xa, va = mirror.lbfgs(ra, x0, UPPER, LOWER, f(x0), 0.001, 100)
    xb, vb = rsched.lbfgs(rb, x0, UPPER, LOWER, f(x0), 0.001, 100)
    assert np.array_equal(ra.candidates(), rb.candidates())
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)) and va == vb
    # the known minimum: the centre clamped to the bounds (uncoupled case),
    # within a few finite-difference steps (tolerance = range * precision)
    tol = (UPPER - LOWER) * 0.001
    if cross == 0.0:
        want = np.clip(center, LOWER, UPPER)
        assert np.all(np.abs(xa - want) <= 20 * tol), (xa, want)
    else:
        # coupled: clamped parameters sit on their bound
        assert xa[0] == UPPER[0] and xa[7] == UPPER[7] and xa[10] == UPPER[10]
    assert va <= f(x0)


def _blob_images():
    from tipl_amd import phantoms
    out = []
    for seed in (0, 1):
        F, M, _ = phantoms.blob_pair(seed, (40, 44, 36))
        out.append((phantoms.quantize(F), phantoms.quantize(M)))
    # masked shapes whose covariance hits each branch of the closed-form
    # eigen solver: axis-aligned box (diagonal), ellipses tilted in the xy,
    # xz and yz planes (one off-diagonal pair), and a general ellipsoid
    z, y, x = np.mgrid[0:32, 0:36, 0:40].astype(np.float64)
    c = np.array([16, 18, 20], np.float64)

    def ell(a, R):
        p = np.stack([z - c[0], y - c[1], x - c[2]], -1) @ R.T
        return ((p / a) ** 2).sum(-1) <= 1

    def rot(i, j, th):
        R = np.eye(3)
        R[i, i] = R[j, j] = np.cos(th)
        R[i, j], R[j, i] = -np.sin(th), np.sin(th)
        return R
    box = np.zeros((32, 36, 40), np.uint8)
    box[6:20, 5:30, 8:35] = 200
    out.append((box, np.roll(box, (2, -1, 3), (0, 1, 2))))
    for R in (rot(1, 2, 0.3), rot(0, 2, -0.4), rot(0, 1, 0.25), rot(0, 1, 0.3) @ rot(1, 2, 0.5)):
        a = np.array([9, 12, 15])
        e1 = (ell(a, R) * 180).astype(np.uint8)
        e2 = (ell(a * [1.1, 0.95, 1.05], R.T) * 150).astype(np.uint8)
        out.append((e1, e2))
    return out


def test_has_mask_and_estimate(mirror, rsched):
    for S, T in _blob_images():
        for v in (S, T):
            assert mirror.has_mask(v) == rsched.has_mask(v)
        for svs, tvs in (((1, 1, 1), (1, 1, 1)), ((1.0, 1.2, 0.9), (2.0, 2.0, 2.0))):
            a = mirror.estimate_affine_param(S, svs, T, tvs)
            b = rsched.estimate_affine_param(S, svs, T, tvs)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (a, b)
    unmasked = np.full((10, 10, 10), 7, np.uint8)
    assert mirror.has_mask(unmasked) == rsched.has_mask(unmasked) is False


def _linear_case(oracle, mirror, rsched, F, M, cost, fvs=(1, 1, 1), tvs=(1, 1, 1), exact=True):
    ca, cb = OracleCosts(oracle), OracleCosts(oracle)
    va, aa, na = mirror.linear(ca, F, fvs, M, tvs, cost)
    vb, ab, nb = rsched.linear(cb, F, fvs, M, tvs, cost)
    assert na == nb == ca.evaluations() == cb.evaluations()
    assert np.float32(va) == np.float32(vb)
    if exact:
        assert np.array_equal(aa.view(np.uint32), ab.view(np.uint32)), (aa, ab)
    else:
        # the masked path ends with transformation_matrix::to_affine_param,
        # whose QR split the mirror computes in double (the reference mixes
        # float and double): the same search, parameters to float rounding
        assert np.all(np.abs(aa - ab) <= 1e-5 * np.maximum(1.0, np.abs(ab))), (aa, ab)
    return aa, na, ca


@pytest.mark.parametrize("cost", ["corr", "mi"])
def test_linear_schedule_small(oracle, mirror, rsched, cost):
    """A 40x44x36 blob pair (no pyramid, no mask): every stage's line search
    and L-BFGS."""
    from tipl_amd import phantoms
    F, M, _ = phantoms.blob_pair(3, (40, 44, 36))
    F, M = phantoms.quantize(F), phantoms.quantize(M)
    assert not (mirror.has_mask(F) and mirror.has_mask(M))
    _linear_case(oracle, mirror, rsched, F, M, cost, (1, 1, 1), (1.1, 1.0, 0.95))


@pytest.mark.parametrize("cost", ["corr", "mi"])
def test_linear_schedule_masked(oracle, mirror, rsched, cost):
    """A masked tissue-phantom pair (zero background): the moment-based start
    (estimate_affine_param), the voxel-size adjustment and the final
    conversion back to the caller's voxel sizes (reg/linear.hpp:652-700)."""
    from tipl_amd import phantoms
    F, M, _ = phantoms.tissue_pair(1001, (44, 52, 44))
    F, M = phantoms.quantize(F), phantoms.quantize(M)
    F[F < 40] = 0                  # a zero background, as after skull stripping
    M[M < 40] = 0
    assert mirror.has_mask(F) and mirror.has_mask(M)
    _linear_case(oracle, mirror, rsched, F, M, cost, (1, 1, 1), (1.2, 1.0, 1.0), exact=False)


def test_linear_schedule_config0(oracle, mirror, rsched):
    """Config 0 at its full 96x112x96 (above 96^3: the half-resolution search,
    then L-BFGS at full size), correlation cost; the recovered affine is
    also compared with the ground truth."""
    from tipl_amd import phantoms
    F, M, truth = phantoms.blob_pair(0)
    F, M = phantoms.quantize(F), phantoms.quantize(M)
    a, n, costs = _linear_case(oracle, mirror, rsched, F, M, "corr")
    assert n > 100
    assert any(k == 1 for k, _ in costs.sets) and any(k > 100 for k, _ in costs.sets)
