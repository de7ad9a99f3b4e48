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


def test_reg_request_defaults():
    """mm_reg's defaults (reg/mm_reg.hpp:53-67): match_fov and masked_r on,
    the reverse direction off unless asked, identity affine for skip_linear."""
    import tipl_amd
    r = tipl_amd.RegRequest()
    tipl_amd.load_library().tiplhip_reg_request_default(ctypes.byref(r))
    assert (r.match_fov, r.masked_r, r.reverse, r.skip_linear) == (1, 1, 0, 0)
    assert list(r.arg) == [0.0] * 6 + [1.0] * 3 + [0.0] * 3
    assert not any([r.f2t, r.to2from, r.from2to, r.warped, r.field])


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors have the C header's sizes and field offsets."""
    import subprocess
    import tipl_amd
    structs = {"tiplhip_reg_request": tipl_amd.RegRequest, "tiplhip_reg_result": tipl_amd.RegResult,
               "tiplhip_cdm_param": tipl_amd.CdmParam, "tiplhip_server_stats": tipl_amd.ServerStats,
               "tiplhip_pair_costs": tipl_amd.PairCosts}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tiplhip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            cfield = fname.rstrip("_")  # `from_` mirrors the C member `from`
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {cfield}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)
