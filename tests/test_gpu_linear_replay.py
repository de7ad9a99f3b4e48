"""linear() on the GPU against the same search replayed on the CPU with the
oracle's costs (tests/cpp/libmirror.so driving the mirror's runner with
oracle costs; the runner itself is pinned to the reference's routines in
test_mirror_host.py).

Correlation costs are exact (integer sums, the same FP64 tail), so the GPU
search must take the same path: the same evaluation count and the same
parameters bit for bit.  MI costs agree to ~1e-12 relative (device log,
summation order), so the bar is the task's: parameters within 1e-3 voxel
(translation) and 1e-3 degree (rotation), scaling and shear within 1e-5."""
import numpy as np
import pytest

from host_shims import Mirror, OracleCosts

pytestmark = pytest.mark.gpu


def _pair(kind, shape):
    from tipl_amd import phantoms
    if kind == "blob":
        F, M, _ = phantoms.blob_pair(0, shape)
        return phantoms.quantize(F), phantoms.quantize(M)
    F, M, _ = phantoms.tissue_pair(1000, shape)
    F, M = phantoms.quantize(F), phantoms.quantize(M)
    F[F < 40] = 0
    M[M < 40] = 0
    return F, M


def _close(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    return (d[0:3].max() <= 1e-3 and np.degrees(d[3:6]).max() <= 1e-3 and d[6:12].max() <= 1e-5)


@pytest.mark.parametrize("kind,shape,cost,bins", [
    ("blob", (96, 112, 96), "corr", 256),       # config 0 at full size
    ("blob", (96, 112, 96), "mi", 256),
    ("tissue", (64, 76, 64), "mi", 64),          # masked: moment start, 64 bins
    ("tissue", (64, 76, 64), "corr", 256),
])
def test_linear_gpu_equals_cpu_replay(ctx, oracle, kind, shape, cost, bins):
    import tipl_amd
    F, M = _pair(kind, shape)
    print(f"case {kind} {shape} {cost} bins={bins}", flush=True)
    g = tipl_amd.register(ctx, F, M, cost=cost, bins=bins)
    assert g["status"] == 0
    costs = OracleCosts(oracle)
    v, a, n = Mirror().linear(costs, F, (1, 1, 1), M, (1, 1, 1), cost, bins=bins)
    print(f"gpu evals {g['evaluations']} cost {g['cost']:.6f}; cpu evals {n} cost {v:.6f}")
    print("gpu", g["affine"], "\ncpu", a)
    if cost == "corr":
        assert g["evaluations"] == n
        assert np.array_equal(g["affine"].view(np.uint32), a.view(np.uint32))
        assert np.float32(g["cost"]) == np.float32(v)
    else:
        assert _close(g["affine"], a), (g["affine"], a)
        assert abs(g["cost"] - v) <= 1e-5 * abs(v)
