"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/tiplhip.h declares, and its host geometry matches the
oracle bit for bit.  No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "tiplhip.h")).read()
    return sorted(set(re.findall(r"\b(tiplhip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import tipl_amd
    lib = tipl_amd.load_library()
    declared = _declared()
    assert len(declared) >= 30
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(tipl_amd.EXPORTS) == declared


def test_affine_matrix_matches_oracle(oracle):
    import tipl_amd
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = np.zeros(12, np.float32)
        p[:3] = rng.uniform(-20, 20, 3)
        p[3:6] = rng.uniform(-1.5, 1.5, 3)
        p[6:9] = rng.uniform(0.5, 2.0, 3)
        p[9:] = rng.uniform(-0.3, 0.3, 3)
        fshape = tuple(int(v) for v in rng.integers(5, 300, 3))
        tshape = tuple(int(v) for v in rng.integers(5, 300, 3))
        fvs = rng.uniform(0.5, 3, 3).astype(np.float32)
        tvs = rng.choice([1.0, 0.8, 2.0], 3).astype(np.float32)
        got = tipl_amd.affine_matrix(p, fshape, fvs, tshape, tvs)
        want = oracle.affine_matrix(p, np.array(fshape[::-1], np.int32), fvs,
                                    np.array(tshape[::-1], np.int32), tvs)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        inv = tipl_amd.matrix_inverse(got)
        inv_o, ok = oracle.matrix_inverse(want)
        assert ok and np.array_equal(inv.view(np.uint32), inv_o.view(np.uint32))


def test_singular_inverse_is_an_error():
    import tipl_amd
    with pytest.raises(tipl_amd.TiplHipError):
        tipl_amd.matrix_inverse(np.zeros(12, np.float32))


def test_errors_are_reported_without_gpu():
    import ctypes
    import tipl_amd
    lib = tipl_amd.load_library()
    out = ctypes.c_void_p()
    rc = lib.tiplhip_ctx_create(-1, ctypes.byref(out))
    assert rc != tipl_amd.OK
    assert lib.tiplhip_last_error()


def test_cdm_param_defaults():
    """cdm_param (reg/cdm.hpp:36-44) defaults, the 200-iteration cap, and the
    fixed-schedule extension off."""
    import tipl_amd
    p = tipl_amd.CdmParam()
    tipl_amd.load_library().tiplhip_cdm_param_default(ctypes.byref(p))
    assert (p.resolution, p.speed, p.smoothing, p.min_dimension, p.max_iterations) == \
        (2.0, np.float32(0.3), np.float32(0.05), 16, 200)
    assert p.gradient_type == b"rrrrrrrr"
    assert (p.levels, p.level_iterations, p.field_bandwidth) == (0, 0, 0)
