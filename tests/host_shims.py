"""ctypes bindings of the host-logic shims used by the optimiser and linear()
pins: tests/cpp/libmirror.so (the C++ mirror's optimisers and runner,
tests/cpp/mirror_shim.cpp) and the sequential-schedule entry points of the
compiled reference (oracle/ref_shim.cpp).  Both are driven by the same cost
callbacks, so their candidate sets and results can be compared directly.
Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MIRROR_SO = os.path.join(ROOT, "tests", "cpp", "libmirror.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_ip = C.POINTER(C.c_int)

COST_CB = C.CFUNCTYPE(None, C.c_void_p, _ip, C.c_void_p, _ip, C.POINTER(C.c_float), C.c_int,
                      C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p)
DOWN_CB = C.CFUNCTYPE(None, C.c_void_p, _ip, C.c_void_p, _ip, C.c_void_p)
PARAM_CB = C.CFUNCTYPE(None, C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_double),
                       C.c_void_p)


def _dims(v):
    return np.array(v.shape[::-1], np.int32)


def _image(ptr, dim_ptr):
    d = [dim_ptr[i] for i in range(3)]
    n = d[0] * d[1] * d[2]
    return np.ctypeslib.as_array((C.c_uint8 * n).from_address(ptr)).reshape(d[2], d[1], d[0])


class OracleCosts:
    """Cost and pyramid callbacks backed by the oracle: every candidate
    transform of a set is priced by oracle.corr_cost / mi_cost (in a small
    thread pool; the oracle's C calls release the GIL), images are halved by
    oracle.downsample.  `sets` records (K, kind) of every call."""

    def __init__(self, oracle, threads=8):
        self.oracle = oracle
        self.pool = ThreadPoolExecutor(threads)
        self.sets = []

        def cost(frm, fdim, to, tdim, T, K, kind, bins, out, user):
            F, G = _image(frm, fdim), _image(to, tdim)
            Ts = np.ctypeslib.as_array(T, (K * 12,)).reshape(K, 12).copy()
            if kind == 1:
                f = lambda t: oracle.mi_cost(F, G, t, bins)
            else:
                f = lambda t: oracle.corr_cost(F, G, t)
            vals = list(self.pool.map(f, Ts)) if K > 1 else [f(Ts[0])]
            for k, v in enumerate(vals):
                out[k] = v
            self.sets.append((K, kind))

        def down(src, dim, dst, odim, user):
            v = _image(src, dim)
            o = oracle.downsample(v)
            n = o.size
            np.ctypeslib.as_array((C.c_uint8 * n).from_address(dst))[:] = o.ravel()
            for i in range(3):
                odim[i] = o.shape[2 - i]

        self.cost_cb = COST_CB(cost)
        self.down_cb = DOWN_CB(down)

    def evaluations(self):
        return sum(k for k, _ in self.sets)


class Recorder:
    """A function of parameter vectors that records every candidate."""

    def __init__(self, f):
        self.f = f
        self.calls = []

        def cb(x, n, K, out, user):
            X = np.ctypeslib.as_array(x, (n * K,)).reshape(K, n).copy()
            self.calls.append(X)
            for k in range(K):
                out[k] = f(X[k])

        self.cb = PARAM_CB(cb)

    def candidates(self):
        return np.concatenate(self.calls) if self.calls else np.zeros((0, 0), np.float32)


class Mirror:
    def __init__(self, path=MIRROR_SO):
        if not os.path.exists(path):
            import pytest
            pytest.skip("tests/cpp/libmirror.so not built (run build())")
        self.lib = C.CDLL(path)
        L = self.lib
        L.mir_linear.restype = C.c_float
        L.mir_linear.argtypes = [_u8p, _i32p, _f32p, _u8p, _i32p, _f32p, C.c_int, C.c_int, C.c_int,
                                 _f32p, COST_CB, DOWN_CB, C.c_void_p, C.POINTER(C.c_uint64)]
        L.mir_has_mask.restype = C.c_int
        L.mir_has_mask.argtypes = [_u8p, _i32p]
        L.mir_estimate_affine_param.restype = None
        L.mir_estimate_affine_param.argtypes = [_u8p, _i32p, _f32p, _u8p, _i32p, _f32p, _f32p]
        L.mir_line_search.restype = C.c_double
        L.mir_line_search.argtypes = [_f32p, C.c_int, _f32p, _f32p, _i32p, C.c_int, C.c_double,
                                      PARAM_CB, C.c_void_p]
        L.mir_lbfgs.restype = C.c_double
        L.mir_lbfgs.argtypes = [_f32p, C.c_int, _f32p, _f32p, C.c_double, C.c_double, C.c_int,
                                PARAM_CB, C.c_void_p]

    def linear(self, costs: OracleCosts, frm, fvs, to, tvs, cost="corr", reg_type=15, bins=256,
               arg=None):
        frm, to = np.ascontiguousarray(frm, np.uint8), np.ascontiguousarray(to, np.uint8)
        a = np.array(arg if arg is not None else [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
        n = C.c_uint64(0)
        r = self.lib.mir_linear(frm, _dims(frm), np.asarray(fvs, np.float32), to, _dims(to),
                                np.asarray(tvs, np.float32), 1 if cost == "mi" else 0, reg_type,
                                bins, a, costs.cost_cb, costs.down_cb, None, C.byref(n))
        return float(r), a, int(n.value)

    def has_mask(self, v):
        v = np.ascontiguousarray(v, np.uint8)
        return bool(self.lib.mir_has_mask(v, _dims(v)))

    def estimate_affine_param(self, S, svs, T, tvs):
        S, T = np.ascontiguousarray(S, np.uint8), np.ascontiguousarray(T, np.uint8)
        a = np.zeros(12, np.float32)
        self.lib.mir_estimate_affine_param(S, _dims(S), np.asarray(svs, np.float32), T, _dims(T),
                                           np.asarray(tvs, np.float32), a)
        return a

    def line_search(self, rec, x, upper, lower, strategy, grid, optimal):
        x = np.array(x, np.float32)
        v = self.lib.mir_line_search(x, x.size, np.asarray(upper, np.float32),
                                     np.asarray(lower, np.float32), np.asarray(strategy, np.int32),
                                     grid, optimal, rec.cb, None)
        return x, v

    def lbfgs(self, rec, x, upper, lower, fun_x, precision=0.001, max_iteration=100):
        x = np.array(x, np.float32)
        v = self.lib.mir_lbfgs(x, x.size, np.asarray(upper, np.float32),
                               np.asarray(lower, np.float32), fun_x, precision, max_iteration,
                               rec.cb, None)
        return x, v


class ReferenceSchedule:
    """The reference's own optimisers and runner (oracle/_ref) driven by the
    same callbacks, with line search and L-BFGS run one after the other."""

    def __init__(self, ref):
        L = ref.lib
        self.lib = L
        L.tref_linear_sequential.restype = C.c_float
        L.tref_linear_sequential.argtypes = [_u8p, _i32p, _f32p, _u8p, _i32p, _f32p, C.c_int,
                                             C.c_int, _f32p, COST_CB, C.c_void_p,
                                             C.POINTER(C.c_uint)]
        L.tref_has_mask.restype = C.c_int
        L.tref_has_mask.argtypes = [_u8p, _i32p]
        L.tref_estimate_affine_param.restype = None
        L.tref_estimate_affine_param.argtypes = [_u8p, _i32p, _f32p, _u8p, _i32p, _f32p, _f32p]
        L.tref_line_search.restype = C.c_double
        L.tref_line_search.argtypes = [_f32p, C.c_int, _f32p, _f32p, _i32p, C.c_int, C.c_double,
                                       PARAM_CB, C.c_void_p]
        L.tref_lbfgs.restype = C.c_double
        L.tref_lbfgs.argtypes = [_f32p, C.c_int, _f32p, _f32p, C.c_double, C.c_double, C.c_int,
                                 PARAM_CB, C.c_void_p]

    def linear(self, costs: OracleCosts, frm, fvs, to, tvs, cost="corr", reg_type=15, arg=None):
        frm, to = np.ascontiguousarray(frm, np.uint8), np.ascontiguousarray(to, np.uint8)
        a = np.array(arg if arg is not None else [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
        n = C.c_uint(0)
        r = self.lib.tref_linear_sequential(frm, _dims(frm), np.asarray(fvs, np.float32), to,
                                            _dims(to), np.asarray(tvs, np.float32),
                                            1 if cost == "mi" else 0, reg_type, a, costs.cost_cb,
                                            None, C.byref(n))
        return float(r), a, int(n.value)

    def has_mask(self, v):
        v = np.ascontiguousarray(v, np.uint8)
        return bool(self.lib.tref_has_mask(v, _dims(v)))

    def estimate_affine_param(self, S, svs, T, tvs):
        S, T = np.ascontiguousarray(S, np.uint8), np.ascontiguousarray(T, np.uint8)
        a = np.zeros(12, np.float32)
        self.lib.tref_estimate_affine_param(S, _dims(S), np.asarray(svs, np.float32), T, _dims(T),
                                            np.asarray(tvs, np.float32), a)
        return a

    def line_search(self, rec, x, upper, lower, strategy, grid, optimal):
        x = np.array(x, np.float32)
        v = self.lib.tref_line_search(x, x.size, np.asarray(upper, np.float32),
                                      np.asarray(lower, np.float32), np.asarray(strategy, np.int32),
                                      grid, optimal, rec.cb, None)
        return x, v

    def lbfgs(self, rec, x, upper, lower, fun_x, precision=0.001, max_iteration=100):
        x = np.array(x, np.float32)
        v = self.lib.tref_lbfgs(x, x.size, np.asarray(upper, np.float32),
                                np.asarray(lower, np.float32), fun_x, precision, max_iteration,
                                rec.cb, None)
        return x, v
