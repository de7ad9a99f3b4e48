"""The reference-side plug (tipl_amd/plug/tipl_hip_plug.cpp): the reference's
own host code -- tipl::reg::linear<tipl::out> and
tipl::reg::cdm_common<tipl::out>, compiled with TIPL_USE_CUDA by g++ -- with
its two GPU plugs (linear_reg_imp<out, mr, uint8, 3>(std::true_type, ...),
reg/linear.hpp:606-609; cdm_cuda<out, uint8, dim>, reg/cdm.hpp:525-531)
defined over libtiplhip.  oracle/_ref/libtiplplug.so holds the plug and a
driver (tests/cpp/plug_driver.cpp); it is built where the reference sources
are present.

CPU: the plug defines the four symbols the reference's CUDA translation units
would (reg/linear_cu.hpp:7-16, reg/cdm_cu.hpp:6-17).
GPU: every cost the reference's optimisers get through the plug equals the
oracle's cost of that transform (bit for bit for correlation, 1e-9 relative
for MI), for the mr = true and mr = false plugs; cdm_common through the plug
gives a field within 0.01 voxel RMS of the oracle's cdm(), in 3-D and in 2-D
(where it is also compared with the reference's own CPU cdm())."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUG_SO = os.path.join(ROOT, "oracle", "_ref", "libtiplplug.so")


@pytest.fixture(scope="module")
def plug():
    if not os.path.exists(PLUG_SO):
        pytest.skip("oracle/_ref/libtiplplug.so not built (needs the reference sources)")
    lib = C.CDLL(PLUG_SO)
    P, vp, ip = C.POINTER, C.c_void_p, C.c_int
    lib.plug_linear.restype = C.c_int64
    lib.plug_linear.argtypes = [vp, P(ip), vp, P(ip), ip, ip, vp, P(C.c_float), ip, vp, vp, P(ip)]
    lib.plug_cdm.restype = ip
    lib.plug_cdm.argtypes = [vp, vp, P(ip), ip, C.c_float, C.c_float, C.c_float, ip, vp]
    lib.ref_cdm2.restype = ip
    lib.ref_cdm2.argtypes = [vp, vp, P(ip), vp]
    return lib


def _dims(shape):
    return (C.c_int * len(shape))(*shape[::-1])


def test_plug_defines_the_reference_gpu_symbols(plug):
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-C", "--defined-only", PLUG_SO], capture_output=True,
                         text=True).stdout
    want = [
        "tipl::reg::linear_reg_imp<tipl::output<0>, true, unsigned char, 3>(std::integral_constant<bool, true>",
        "tipl::reg::linear_reg_imp<tipl::output<0>, false, unsigned char, 3>(std::integral_constant<bool, true>",
        "tipl::reg::cdm_cuda<tipl::output<0>, unsigned char, 3>(",
        "tipl::reg::cdm_cuda<tipl::output<0>, unsigned char, 2>(",
    ]
    lines = [l for l in out.splitlines() if " T " in l]
    for w in want:
        assert any(w in l for l in lines), w


@pytest.mark.gpu
def test_reference_cdm_common_2d_through_plug(plug, oracle):
    """cdm_common<out, uint8, 2> (cuda = true) reaches cdm_cuda<out, uint8, 2>,
    which the plug runs on the device through tiplhip_cdm2d."""
    from scipy.ndimage import gaussian_filter, shift
    rng = np.random.default_rng(5)
    a = gaussian_filter(rng.random((96, 112)), 3.0)
    a = ((a - a.min()) / (a.max() - a.min()) * 255.99).astype(np.uint8)
    b = np.clip(shift(a.astype(np.float32), (1.3, -0.8), order=1), 0, 255).astype(np.uint8)
    f_plug = np.zeros(a.shape + (2,), np.float32)
    f_ref = np.zeros(a.shape + (2,), np.float32)
    print("reference cdm_common (2-D) through the plug", flush=True)
    assert plug.plug_cdm(a.ctypes.data, b.ctypes.data, _dims(a.shape), 2, 2.0, 0.3, 0.05, 16,
                         f_plug.ctypes.data) == 1
    ok, fo, _ = oracle.cdm2(a, b)
    assert ok
    assert np.abs(f_plug).max() > 0.1
    assert float(np.sqrt(((f_plug - fo) ** 2).sum(-1).mean())) <= 0.01
    # and the reference's own CPU cdm() (its threads sum the cost in varying
    # order, so the fields agree to rounding)
    assert plug.ref_cdm2(a.ctypes.data, b.ctypes.data, _dims(a.shape), f_ref.ctypes.data) == 1
    assert float(np.sqrt(((f_plug - f_ref) ** 2).sum(-1).mean())) <= 0.01


def _blob_pair(shape=(40, 48, 40)):
    from tipl_amd import phantoms
    F, M, _ = phantoms.blob_pair(3, shape)
    return phantoms.quantize(F), phantoms.quantize(M)


@pytest.mark.gpu
@pytest.mark.parametrize("mi,narrow", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_reference_linear_through_plug(plug, oracle, mi, narrow):
    F, M = _blob_pair()
    arg = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
    cap = 4000
    logT = np.zeros((cap, 12), np.float32)
    logc = np.zeros(cap, np.float64)
    n = C.c_int(0)
    cost = C.c_float(0)
    print(f"reference linear() through the plug: mi={mi} narrow={narrow}", flush=True)
    calls = plug.plug_linear(F.ctypes.data, _dims(F.shape), M.ctypes.data, _dims(M.shape), mi,
                             narrow, arg.ctypes.data, C.byref(cost), cap, logT.ctypes.data,
                             logc.ctypes.data, C.byref(n))
    print(f"{calls} plug calls, final cost {cost.value}, arg {arg}")
    assert calls > 20 and n.value == min(calls, cap)
    assert np.isfinite(arg).all()
    idx = np.unique(np.linspace(0, n.value - 1, 60).astype(int))
    for i in idx:
        if mi:
            want = oracle.mi_cost(F, M, logT[i])
            assert abs(logc[i] - want) <= 1e-9 * abs(want), i
        else:
            assert logc[i] == oracle.corr_cost(F, M, logT[i]), i
    # the optimiser ends no worse than where it started
    start = (oracle.mi_cost if mi else oracle.corr_cost)(F, M, logT[0])
    assert logc[:n.value].min() <= start


@pytest.mark.gpu
def test_reference_cdm_common_through_plug(plug, oracle):
    from tipl_amd import phantoms
    F, M, _ = phantoms.warped_pair(2100, (40, 48, 40))
    It, Is = phantoms.quantize(F), phantoms.quantize(M)
    field = np.zeros(It.shape + (3,), np.float32)
    print("reference cdm_common through the plug", flush=True)
    assert plug.plug_cdm(It.ctypes.data, Is.ctypes.data, _dims(It.shape), 3, 2.0, 0.3, 0.05, 16,
                         field.ctypes.data) == 1
    ok, fo, _ = oracle.cdm(It, Is)
    assert ok
    assert np.abs(fo).max() > 0.1
    rms = float(np.sqrt(((field - fo) ** 2).sum(-1).mean()))
    assert rms <= 0.01, rms
