"""ctypes bindings for the checker libraries (test infrastructure only).

``Oracle`` wraps liboracle.so, the C restatement of TIPL's registration hot
path (oracle/oracle.c).  ``Reference`` wraps oracle/_ref/libtiplref.so, the
reference headers compiled through oracle/ref_shim.cpp; it exists only where
the reference sources were available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product (tipl_amd, libtiplhip.so) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtiplref.so")

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")


def _dims(vol: np.ndarray) -> np.ndarray:
    """x-fastest dims {w,h,d} of a C-ordered (z,y,x) array."""
    return np.array(vol.shape[::-1], dtype=np.int32)


def _f3(field: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(field, dtype=np.float32)


class _Lib:
    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)

    def _fn(self, name, res, *args):
        f = getattr(self.lib, name)
        f.restype = res
        f.argtypes = list(args)
        return f


class _Warp:
    """Post-registration warping (numerical/dif.hpp) and mm_reg's r, shared
    by both libraries: the C entry points have the same signatures."""

    def _bind_warp(self, prefix: str):
        _vp = C.c_void_p
        self._d2m = self._fn(prefix + "disp_to_mapping", None, _f32p, _i32p, _vp, _f32p)
        self._m2d = (self._fn(prefix + "mapping_to_disp", None, _f32p, _i32p, _f32p)
                     if hasattr(self.lib, prefix + "mapping_to_disp") else None)
        self._inv_d2m = self._fn(prefix + "inv_disp_to_mapping", None, _f32p, _i32p, _f32p, _i32p,
                                 _f32p)
        self._cm_u8 = self._fn(prefix + "compose_mapping_u8", None, _u8p, _i32p, _f32p, _i32p, _u8p)
        self._cm_f32 = self._fn(prefix + "compose_mapping_f32", None, _f32p, _i32p, _f32p, _i32p,
                                _f32p)
        self._jdd = self._fn(prefix + "jacobian_det_dis", None, _f32p, _i32p, _f32p)
        self._jd = self._fn(prefix + "jacobian_det", None, _f32p, _i32p, _f32p)
        self._corr_r = self._fn(prefix + "correlation", C.c_double, _u8p, _u8p, C.c_size_t, C.c_int)

    @staticmethod
    def _fdims(field):
        return np.array(field.shape[:3][::-1], np.int32)

    def displacement_to_mapping(self, dis, T=None) -> np.ndarray:
        dis = _f3(dis)
        out = np.zeros_like(dis)
        Tp = None
        if T is not None:
            Tarr = np.ascontiguousarray(T, np.float32)
            Tp = Tarr.ctypes.data
        self._d2m(dis, self._fdims(dis), Tp, out)
        return out

    def mapping_to_displacement(self, mapping) -> np.ndarray:
        mapping = _f3(mapping)
        out = np.zeros_like(mapping)
        self._m2d(mapping, self._fdims(mapping), out)
        return out

    def inv_displacement_to_mapping(self, inv_dis, T, out_shape) -> np.ndarray:
        inv_dis = _f3(inv_dis)
        out = np.zeros(tuple(out_shape) + (3,), np.float32)
        self._inv_d2m(inv_dis, self._fdims(inv_dis), np.ascontiguousarray(T, np.float32),
                      np.array(tuple(out_shape)[::-1], np.int32), out)
        return out

    def compose_mapping(self, src, mapping) -> np.ndarray:
        mapping = _f3(mapping)
        shape = mapping.shape[:3]
        if np.asarray(src).dtype == np.uint8:
            src = np.ascontiguousarray(src, np.uint8)
            out = np.zeros(shape, np.uint8)
            self._cm_u8(src, _dims(src), mapping, self._fdims(mapping), out)
        else:
            src = np.ascontiguousarray(src, np.float32)
            out = np.zeros(shape, np.float32)
            self._cm_f32(src, _dims(src), mapping, self._fdims(mapping), out)
        return out

    def jacobian_determinant_dis(self, dis) -> np.ndarray:
        dis = _f3(dis)
        out = np.zeros(dis.shape[:3], np.float32)
        self._jdd(dis, self._fdims(dis), out)
        return out

    def jacobian_determinant(self, mapping) -> np.ndarray:
        mapping = _f3(mapping)
        out = np.zeros(mapping.shape[:3], np.float32)
        self._jd(mapping, self._fdims(mapping), out)
        return out

    def correlation(self, x, y, masked=True) -> float:
        x = np.ascontiguousarray(x, np.uint8).ravel()
        y = np.ascontiguousarray(y, np.uint8).ravel()
        assert x.size == y.size
        return float(self._corr_r(x, y, x.size, int(bool(masked))))


class _Cdm2:
    """The 2-D CDM pieces (cdm<out,uint8,2>), shared by both libraries.
    Images are C-ordered (h, w) uint8 arrays; fields (h, w, 2) float32."""

    def _bind_cdm2(self, prefix: str):
        P = C.POINTER
        self._c2_prefix = prefix
        self._c2_sample = self._fn(prefix + "sample_u8", C.c_uint8, _u8p, _i32p, C.c_float, C.c_float)
        self._c2_down = self._fn(prefix + "downsample_u8", None, _u8p, _i32p, _u8p, _i32p)
        self._c2_up = self._fn(prefix + "upsample_f2", None, _f32p, _i32p, _f32p, _i32p)
        self._c2_compose = self._fn(prefix + "compose_displacement", None, _u8p, _i32p, _f32p, _u8p)
        self._c2_grad = self._fn(prefix + "cdm_gradient", C.c_float, _u8p, _u8p, _i32p, C.c_char, _f32p)
        self._c2_acc = self._fn(prefix + "accumulate_displacement", None, _f32p, _f32p, _i32p)
        self._c2_smooth = self._fn(prefix + "cdm_smooth", None, _f32p, _f32p, _i32p, C.c_float)
        if prefix == "orc2_":
            self._c2_poisson = self._fn("orc2_cdm_poisson", None, _f32p, _i32p, C.c_int)
            self._c2_maxlen = self._fn("orc2_cdm_max_length", C.c_float, _f32p, C.c_size_t)
            self._c2_cdm = self._fn("orc2_cdm", C.c_int, P(C.c_void_p), P(C.c_void_p), C.c_int,
                                    _i32p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                    C.c_char, _f32p, P(C.c_int))
        else:
            self._c2_poisson = self._fn("tref2_cdm_poisson", None, _f32p, _i32p)
            self._c2_maxlen = self._fn("tref2_cdm_max_length", C.c_float, _f32p, _i32p)
            self._c2_cdm = self._fn("tref2_cdm", C.c_int, P(C.c_void_p), P(C.c_void_p), C.c_int,
                                    _i32p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_char,
                                    _f32p)

    @staticmethod
    def _d2(a) -> np.ndarray:
        """x-fastest dims {w, h} of a (h, w) image or (h, w, 2) field."""
        return np.array([a.shape[1], a.shape[0]], np.int32)

    def sample2(self, img, x, y) -> int:
        img = np.ascontiguousarray(img, np.uint8)
        return int(self._c2_sample(img, self._d2(img), x, y))

    def downsample2(self, v) -> np.ndarray:
        v = np.ascontiguousarray(v, np.uint8)
        out = np.zeros(((v.shape[0] + 1) // 2, (v.shape[1] + 1) // 2), np.uint8)
        od = np.zeros(2, np.int32)
        self._c2_down(v, self._d2(v), out, od)
        return out

    def upsample2(self, field, out_shape) -> np.ndarray:
        field = np.ascontiguousarray(field, np.float32)
        out = np.zeros(tuple(out_shape) + (2,), np.float32)
        self._c2_up(field, self._d2(field), out, np.array(tuple(out_shape)[::-1], np.int32))
        return out

    def compose2(self, Is, d) -> np.ndarray:
        Is = np.ascontiguousarray(Is, np.uint8)
        out = np.zeros_like(Is)
        self._c2_compose(Is, self._d2(Is), np.ascontiguousarray(d, np.float32), out)
        return out

    def gradient2(self, Js, It, kind="r"):
        Js = np.ascontiguousarray(Js, np.uint8)
        It = np.ascontiguousarray(It, np.uint8)
        nd = np.zeros(Js.shape + (2,), np.float32)
        c = self._c2_grad(Js, It, self._d2(Js), kind.encode(), nd)
        return nd, float(c)

    def poisson2(self, nd) -> np.ndarray:
        """The reference's 12 Jacobi sweeps."""
        nd = np.array(nd, np.float32, order="C")
        if self._c2_prefix == "orc2_":
            self._c2_poisson(nd, self._d2(nd), 12)
        else:
            self._c2_poisson(nd, self._d2(nd))
        return nd

    def max_length2(self, d) -> float:
        d = np.ascontiguousarray(d, np.float32)
        if self._c2_prefix == "orc2_":
            return float(self._c2_maxlen(d, d.size // 2))
        return float(self._c2_maxlen(d, self._d2(d)))

    def accumulate2(self, cur, nd) -> np.ndarray:
        cur = np.ascontiguousarray(cur, np.float32)
        nd = np.array(nd, np.float32, order="C")
        self._c2_acc(cur, nd, self._d2(cur))
        return nd

    def smooth2(self, d, dd, s) -> np.ndarray:
        d = np.ascontiguousarray(d, np.float32)
        dd = np.array(dd, np.float32, order="C")
        self._c2_smooth(d, dd, self._d2(d), s)
        return dd

    def cdm2(self, It, Is, resolution=2.0, speed=0.3, smoothing=0.05, min_dimension=16,
             gradient_type="r", max_iterations=200):
        """cdm() on 2-D images; returns (ok, field, iterations or None)."""
        if isinstance(It, np.ndarray):
            It, Is = [It], [Is]
        It = [np.ascontiguousarray(v, np.uint8) for v in It]
        Is = [np.ascontiguousarray(v, np.uint8) for v in Is]
        n = len(It)
        pt = (C.c_void_p * n)(*[v.ctypes.data for v in It])
        ps = (C.c_void_p * n)(*[v.ctypes.data for v in Is])
        field = np.zeros(It[0].shape + (2,), np.float32)
        if self._c2_prefix == "orc2_":
            iters = C.c_int(0)
            ok = self._c2_cdm(pt, ps, n, self._d2(It[0]), resolution, speed, smoothing,
                              min_dimension, max_iterations, gradient_type.encode(), field,
                              C.byref(iters))
            return bool(ok), field, iters.value
        ok = self._c2_cdm(pt, ps, n, self._d2(It[0]), resolution, speed, smoothing, min_dimension,
                          gradient_type.encode(), field)
        return bool(ok), field, None


class Oracle(_Lib, _Warp, _Cdm2):
    """The C restatement; every method mirrors one reference routine."""

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        P = C.POINTER
        self._affine = self._fn("orc_affine_matrix", None, _f32p, _i32p, _f32p, _i32p, _f32p, _f32p)
        self._inv = self._fn("orc_matrix_inverse", C.c_int, _f32p)
        self._resample = self._fn("orc_resample_u8", None, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._resample_nc = self._fn("orc_resample_u8_nocrop", None, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._sample = self._fn("orc_sample_u8", C.c_uint8, _u8p, _i32p, C.c_float, C.c_float, C.c_float)
        self._sums = self._fn("orc_corr_sums", None, _u8p, _u8p, C.c_size_t, _i64p)
        self._corr_from_sums = self._fn("orc_corr_cost_from_sums", C.c_double, _i64p, C.c_size_t)
        self._corr = self._fn("orc_corr_cost", C.c_double, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._mi_hist = self._fn("orc_mi_hist", None, _u8p, _i32p, _u8p, _i32p, _f32p, C.c_int, _i32p)
        self._vhist = self._fn("orc_value_hist", None, _u8p, C.c_size_t, C.c_int, _u32p)
        self._mi_mean = self._fn("orc_mi_mean", C.c_double, _i32p, _u32p, C.c_int)
        self._mi = self._fn("orc_mi_cost", C.c_double, _u8p, _i32p, _u8p, _i32p, _f32p, C.c_int)
        self._norm = self._fn("orc_normalize_f32", C.c_int, _f32p, C.c_size_t, _u8p, C.c_float)
        self._norm16 = self._fn("orc_normalize_i16", C.c_int, _i16p, C.c_size_t, _u8p, C.c_float)
        self._down = self._fn("orc_downsample_u8", None, _u8p, _i32p, _u8p, _i32p)
        self._up = self._fn("orc_upsample_f3", None, _f32p, _i32p, _f32p, _i32p)
        self._compose = self._fn("orc_compose_displacement", None, _u8p, _i32p, _f32p, _u8p)
        self._grad = self._fn("orc_cdm_gradient", C.c_float, _u8p, _u8p, _i32p, C.c_char, _f32p)
        self._poisson = self._fn("orc_cdm_poisson", None, _f32p, _i32p, C.c_int)
        self._maxlen = self._fn("orc_cdm_max_length", C.c_float, _f32p, C.c_size_t)
        self._acc = self._fn("orc_accumulate_displacement", None, _f32p, _f32p, _i32p)
        self._smooth = self._fn("orc_cdm_smooth", None, _f32p, _f32p, _i32p, C.c_float)
        self._improved = self._fn("orc_cdm_improved", C.c_int, _f32p, P(C.c_int))
        self._gauss = self._fn("orc_gaussian_f32", None, _f32p, _i32p, C.c_int, C.c_uint)
        self._subj = self._fn("orc_subject_pre", None, _f32p, _i32p, _u8p)
        self._otsu_median = self._fn("orc_otsu_median", C.c_float, _f32p, C.c_size_t)
        self._median = self._fn("orc_median", C.c_float, _f32p, C.c_size_t)
        self._otsu = self._fn("orc_otsu_threshold", C.c_float, _f32p, C.c_size_t)
        self._cdm_fixed = self._fn("orc_cdm_fixed", None, P(C.c_void_p), P(C.c_void_p), C.c_int,
                                   _i32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint,
                                   C.c_char, _f32p, P(C.c_int))
        self._cdm = self._fn("orc_cdm", C.c_int, P(C.c_void_p), P(C.c_void_p), C.c_int, _i32p,
                             C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_char, _f32p,
                             P(C.c_int))
        self._bind_warp("orc_")
        self._bind_cdm2("orc2_")
        self._fov = self._fn("orc_match_fov", None, _u8p, _i32p, _f32p, _i32p)

    def match_fov(self, It, T, moving_shape) -> np.ndarray:
        """mm_reg::It_match_fov on a copy of It (T maps It voxels to the moving grid)."""
        out = np.array(It, np.uint8, order="C")
        self._fov(out, _dims(out), np.ascontiguousarray(T, np.float32),
                  np.array(tuple(moving_shape)[::-1], np.int32))
        return out

    # -- geometry ---------------------------------------------------------
    def affine_matrix(self, param, fdim, fvs, tdim, tvs) -> np.ndarray:
        T = np.zeros(12, np.float32)
        self._affine(np.asarray(param, np.float32), np.asarray(fdim, np.int32),
                     np.asarray(fvs, np.float32), np.asarray(tdim, np.int32),
                     np.asarray(tvs, np.float32), T)
        return T

    def matrix_inverse(self, T):
        T = np.array(T, np.float32)
        ok = self._inv(T)
        return T, bool(ok)

    # -- resampling ---------------------------------------------------------
    def resample(self, src: np.ndarray, out_shape, T, crop: bool = True) -> np.ndarray:
        src = np.ascontiguousarray(src, np.uint8)
        out = np.zeros(out_shape, np.uint8)
        f = self._resample if crop else self._resample_nc
        f(src, _dims(src), out, _dims(out), np.asarray(T, np.float32))
        return out

    def sample(self, src, x, y, z) -> int:
        src = np.ascontiguousarray(src, np.uint8)
        return int(self._sample(src, _dims(src), x, y, z))

    # -- costs --------------------------------------------------------------
    def corr_sums(self, x, y) -> np.ndarray:
        s = np.zeros(5, np.int64)
        x = np.ascontiguousarray(x, np.uint8).ravel()
        y = np.ascontiguousarray(y, np.uint8).ravel()
        self._sums(x, y, x.size, s)
        return s

    def corr_cost_from_sums(self, s, n) -> float:
        return float(self._corr_from_sums(np.asarray(s, np.int64), n))

    def corr_cost(self, frm, to, T) -> float:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        return float(self._corr(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32)))

    def mi_hist(self, frm, to, T, bins=256) -> np.ndarray:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        h = np.zeros(bins * bins, np.int32)
        self._mi_hist(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32), bins, h)
        return h.reshape(bins, bins)

    def value_hist(self, v, bins=256) -> np.ndarray:
        v = np.ascontiguousarray(v, np.uint8).ravel()
        h = np.zeros(bins, np.uint32)
        self._vhist(v, v.size, bins, h)
        return h

    def mi_mean(self, hist, from_hist) -> float:
        hist = np.ascontiguousarray(hist, np.int32)
        bins = hist.shape[0]
        return float(self._mi_mean(hist.ravel(), np.ascontiguousarray(from_hist, np.uint32), bins))

    def mi_cost(self, frm, to, T, bins=256) -> float:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        return float(self._mi(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32), bins))

    def normalize(self, v, upper=255.99) -> np.ndarray:
        out = np.zeros(v.shape, np.uint8)
        if v.dtype == np.int16:
            self._norm16(np.ascontiguousarray(v).ravel(), v.size, out.reshape(-1), upper)
        else:
            self._norm(np.ascontiguousarray(v, np.float32).ravel(), v.size, out.reshape(-1), upper)
        return out

    # -- pyramid --------------------------------------------------------------
    def downsample(self, v) -> np.ndarray:
        v = np.ascontiguousarray(v, np.uint8)
        od = np.zeros(3, np.int32)
        s = v.shape
        out = np.zeros(((s[0] + 1) // 2, (s[1] + 1) // 2, (s[2] + 1) // 2), np.uint8)
        self._down(v, _dims(v), out, od)
        return out

    def upsample(self, field, out_shape) -> np.ndarray:
        field = _f3(field)
        out = np.zeros(tuple(out_shape) + (3,), np.float32)
        self._up(field, np.array(field.shape[:3][::-1], np.int32), out,
                 np.array(tuple(out_shape)[::-1], np.int32))
        return out

    # -- CDM pieces ------------------------------------------------------------
    def compose(self, Is, d) -> np.ndarray:
        Is = np.ascontiguousarray(Is, np.uint8)
        out = np.zeros_like(Is)
        self._compose(Is, _dims(Is), _f3(d), out)
        return out

    def gradient(self, Js, It, kind="r"):
        Js = np.ascontiguousarray(Js, np.uint8)
        It = np.ascontiguousarray(It, np.uint8)
        nd = np.zeros(Js.shape + (3,), np.float32)
        c = self._grad(Js, It, _dims(Js), kind.encode(), nd)
        return nd, float(c)

    def poisson(self, nd, sweeps=12) -> np.ndarray:
        nd = _f3(nd).copy()
        self._poisson(nd, np.array(nd.shape[:3][::-1], np.int32), sweeps)
        return nd

    def max_length(self, d) -> float:
        d = _f3(d)
        return float(self._maxlen(d, d.size // 3))

    def accumulate(self, cur, nd) -> np.ndarray:
        cur = _f3(cur)
        nd = _f3(nd).copy()
        self._acc(cur, nd, np.array(cur.shape[:3][::-1], np.int32))
        return nd

    def smooth(self, d, dd, s) -> np.ndarray:
        d = _f3(d)
        dd = _f3(dd).copy()
        self._smooth(d, dd, np.array(d.shape[:3][::-1], np.int32), s)
        return dd

    def gaussian(self, img, bandwidth=1) -> np.ndarray:
        """filter::gaussian on a float32 (z, y, x) image or (z, y, x, 3) field; returns a copy."""
        out = np.array(img, np.float32, order="C")
        ch = 3 if out.ndim == 4 else 1
        assert out.ndim == 3 or out.shape[3] == 3, "scalar image or vector<3> field"
        self._gauss(out.reshape(-1), np.array(out.shape[:3][::-1], np.int32), ch, int(bandwidth))
        return out

    def improved(self, cost):
        c = np.array(list(cost) + [0.0] * 2, np.float32)
        n = C.c_int(len(cost))
        r = self._improved(c, C.byref(n))
        return bool(r), list(map(float, c[: n.value]))

    def subject_pre(self, img) -> np.ndarray:
        """reg::subject_image_pre of a float32 (z, y, x) image -> uint8."""
        img = np.ascontiguousarray(img, np.float32)
        out = np.zeros(img.shape, np.uint8)
        self._subj(img.reshape(-1), _dims(img), out.reshape(-1))
        return out

    def otsu_median(self, v) -> float:
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._otsu_median(v, v.size))

    def otsu_threshold(self, v) -> float:
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._otsu(v, v.size))

    def median(self, v) -> float:
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._median(v, v.size))

    def cdm_fixed(self, It, Is, levels=3, level_iterations=48, speed=0.3, smoothing=0.05,
                  bandwidth=8, gradient_type="r"):
        """The fixed-schedule CDM extension; returns (field, iterations)."""
        It = [np.ascontiguousarray(v, np.uint8) for v in (It if isinstance(It, list) else [It])]
        Is = [np.ascontiguousarray(v, np.uint8) for v in (Is if isinstance(Is, list) else [Is])]
        n = len(It)
        pt = (C.c_void_p * n)(*[v.ctypes.data for v in It])
        ps = (C.c_void_p * n)(*[v.ctypes.data for v in Is])
        field = np.zeros(It[0].shape + (3,), np.float32)
        it = C.c_int(0)
        self._cdm_fixed(pt, ps, n, _dims(It[0]), levels, level_iterations, speed, smoothing,
                        bandwidth, gradient_type.encode(), field, C.byref(it))
        return field, it.value

    def cdm(self, It, Is, resolution=2.0, speed=0.3, smoothing=0.05, min_dimension=16,
            max_iterations=200, gradient_type="r"):
        if isinstance(It, np.ndarray):
            It, Is = [It], [Is]
        It = [np.ascontiguousarray(v, np.uint8) for v in It]
        Is = [np.ascontiguousarray(v, np.uint8) for v in Is]
        n = len(It)
        pt = (C.c_void_p * n)(*[v.ctypes.data for v in It])
        ps = (C.c_void_p * n)(*[v.ctypes.data for v in Is])
        field = np.zeros(It[0].shape + (3,), np.float32)
        iters = C.c_int(0)
        ok = self._cdm(pt, ps, n, _dims(It[0]), resolution, speed, smoothing, min_dimension,
                       max_iterations, gradient_type.encode(), field, C.byref(iters))
        return bool(ok), field, iters.value


class Reference(_Lib, _Warp, _Cdm2):
    """The reference headers compiled through oracle/ref_shim.cpp."""

    def __init__(self, path: str = REF_SO):
        super().__init__(path)
        P = C.POINTER
        self._threads = self._fn("tref_set_threads", None, C.c_int)
        self._get_threads = self._fn("tref_get_threads", C.c_int)
        self._affine = self._fn("tref_affine_matrix", None, _f32p, _i32p, _f32p, _i32p, _f32p, _f32p)
        self._inv = self._fn("tref_matrix_inverse", C.c_int, _f32p)
        self._resample = self._fn("tref_resample_u8", None, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._sample = self._fn("tref_sample_u8", C.c_uint8, _u8p, _i32p, C.c_float, C.c_float, C.c_float)
        self._corr = self._fn("tref_corr_cost", C.c_double, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._mi_hist = self._fn("tref_mi_hist", None, _u8p, _i32p, _u8p, _i32p, _f32p, _i32p)
        self._mi = self._fn("tref_mi_cost", C.c_double, _u8p, _i32p, _u8p, _i32p, _f32p)
        self._batch = self._fn("tref_cost_batch", None, C.c_int, _u8p, _i32p, _u8p, _i32p, _f32p,
                               C.c_int, _f64p)
        self._norm = self._fn("tref_normalize_f32", C.c_int, _f32p, C.c_size_t, _i32p, _u8p, C.c_float)
        self._down = self._fn("tref_downsample_u8", None, _u8p, _i32p, _u8p, _i32p)
        self._up = self._fn("tref_upsample_f3", None, _f32p, _i32p, _f32p, _i32p)
        self._compose = self._fn("tref_compose_displacement", None, _u8p, _i32p, _f32p, _u8p)
        self._grad = self._fn("tref_cdm_gradient", C.c_float, _u8p, _u8p, _i32p, C.c_char, _f32p)
        self._poisson = self._fn("tref_cdm_poisson", None, _f32p, _i32p)
        self._maxlen = self._fn("tref_cdm_max_length", C.c_float, _f32p, _i32p)
        self._acc = self._fn("tref_accumulate_displacement", None, _f32p, _f32p, _i32p)
        self._smooth = self._fn("tref_cdm_smooth", None, _f32p, _f32p, _i32p, C.c_float)
        self._improved = self._fn("tref_cdm_improved", C.c_int, _f32p, P(C.c_int))
        self._cdm = self._fn("tref_cdm", C.c_int, P(C.c_void_p), P(C.c_void_p), C.c_int, _i32p,
                             C.c_float, C.c_float, C.c_float, C.c_int, C.c_char, _f32p)
        self._linear = self._fn("tref_linear", C.c_float, _u8p, _i32p, _f32p, _u8p, _i32p, _f32p,
                                C.c_int, C.c_int, _f32p)
        self._gauss = self._fn("tref_gaussian_f32", None, _f32p, _i32p, C.c_int, C.c_uint)
        self._cdm_fixed = self._fn("tref_cdm_fixed", None, _u8p, _u8p, _i32p, C.c_int, C.c_int,
                                   C.c_float, C.c_uint, _f32p)
        self._subj = self._fn("tref_subject_pre", None, _f32p, _i32p, _u8p)
        self._otsu_median = self._fn("tref_otsu_median", C.c_float, _f32p, C.c_size_t)
        self._median = self._fn("tref_median", C.c_float, _f32p, C.c_size_t)
        self._otsu = self._fn("tref_otsu_threshold", C.c_float, _f32p, C.c_size_t)
        self._norm16 = self._fn("tref_normalize_i16", C.c_int, _i16p, C.c_size_t, _i32p, _u8p,
                                C.c_float)
        self._cdm_iters = self._fn("tref_cdm_iterations", C.c_double, _u8p, _u8p, _i32p, C.c_int,
                                   C.c_void_p)
        self._bind_warp("tref_")
        self._bind_cdm2("tref2_")
        self._fov = self._fn("tref_match_fov", None, _u8p, _i32p, _f32p, _i32p, _f32p, _f32p)
        _ip = C.POINTER(C.c_int)
        self._nii_f = self._fn("tref_nifti_read_f32", C.c_int, C.c_char_p, C.c_int, _f32p,
                               C.c_size_t, _i32p, _f32p, _f32p, _ip)
        self._nii_u8 = self._fn("tref_nifti_read_u8", C.c_int, C.c_char_p, C.c_int, _u8p,
                                C.c_size_t, _i32p, _f32p, _f32p, _ip)
        self._nii_w = self._fn("tref_nifti_write_f32", C.c_int, C.c_char_p, _f32p, _i32p, _f32p,
                               _f32p)
        self._mm = self._fn("tref_mm_reg", None, _u8p, _i32p, _f32p, _u8p, _i32p, _f32p, _f32p,
                            C.c_int, C.c_int, C.c_int, _f32p, _u8p, _u8p, _f32p, _f32p, _f32p,
                            _f32p, _u8p)

    def match_fov(self, It, tvs, moving_shape, mvs, arg) -> np.ndarray:
        out = np.array(It, np.uint8, order="C")
        self._fov(out, _dims(out), np.asarray(tvs, np.float32),
                  np.array(tuple(moving_shape)[::-1], np.int32), np.asarray(mvs, np.float32),
                  np.asarray(arg, np.float32))
        return out

    def mm_reg(self, It, tvs, I, svs, arg, reverse=True, match_fov=True, masked_r=True):
        """The reference mm_reg's nonlinear half with a given affine; returns a dict."""
        It = np.ascontiguousarray(It, np.uint8)
        I = np.ascontiguousarray(I, np.uint8)
        f = lambda shp: np.zeros(tuple(shp) + (3,), np.float32)
        out = dict(r=np.zeros(2, np.float32), J_lin=np.zeros_like(It), It_fov=np.zeros_like(It),
                   t2f=f(It.shape), f2t=f(It.shape), to2from=f(It.shape), from2to=f(I.shape),
                   JJ=np.zeros_like(It))
        self._mm(It, _dims(It), np.asarray(tvs, np.float32), I, _dims(I),
                 np.asarray(svs, np.float32), np.asarray(arg, np.float32), int(reverse),
                 int(match_fov), int(masked_r), out["r"], out["J_lin"], out["It_fov"], out["t2f"],
                 out["f2t"], out["to2from"], out["from2to"], out["JJ"])
        return out

    def nifti_read(self, path, volume=0, dtype=np.float32, cap=1 << 26):
        """tipl::io::gz_nifti `>>` of one volume: (array (z, y, x), vs, srow, is_mni)."""
        buf = np.zeros(cap, np.dtype(dtype))
        dim = np.zeros(3, np.int32)
        vs = np.zeros(3, np.float32)
        srow = np.zeros(12, np.float32)
        mni = C.c_int(0)
        f = self._nii_u8 if np.dtype(dtype) == np.uint8 else self._nii_f
        ok = f(os.fsencode(str(path)), int(volume), buf, cap, dim, vs, srow, C.byref(mni))
        if not ok:
            raise IOError(f"reference reader failed on {path}")
        n = int(np.prod(dim))
        return buf[:n].reshape(dim[::-1]).copy(), vs, srow, bool(mni.value)

    def nifti_write(self, path, data, vs, srow):
        data = np.ascontiguousarray(data, np.float32)
        ok = self._nii_w(os.fsencode(str(path)), data.reshape(-1), _dims(data),
                         np.asarray(vs, np.float32), np.asarray(srow, np.float32))
        if not ok:
            raise IOError(f"reference writer failed on {path}")

    def cdm_iterations(self, It, Is, iters, want_field=False):
        """Wall seconds of `iters` reference cdm() loop-body iterations at one level."""
        It = np.ascontiguousarray(It, np.uint8)
        Is = np.ascontiguousarray(Is, np.uint8)
        field = np.zeros(It.shape + (3,), np.float32) if want_field else None
        sec = self._cdm_iters(It, Is, _dims(It), int(iters),
                              field.ctypes.data if want_field else None)
        return (float(sec), field) if want_field else float(sec)

    def set_threads(self, n: int):
        self._threads(int(n))

    def threads(self) -> int:
        return int(self._get_threads())

    def affine_matrix(self, param, fdim, fvs, tdim, tvs):
        T = np.zeros(12, np.float32)
        self._affine(np.asarray(param, np.float32), np.asarray(fdim, np.int32),
                     np.asarray(fvs, np.float32), np.asarray(tdim, np.int32),
                     np.asarray(tvs, np.float32), T)
        return T

    def matrix_inverse(self, T):
        T = np.array(T, np.float32)
        ok = self._inv(T)
        return T, bool(ok)

    def resample(self, src, out_shape, T):
        src = np.ascontiguousarray(src, np.uint8)
        out = np.zeros(out_shape, np.uint8)
        self._resample(src, _dims(src), out, _dims(out), np.asarray(T, np.float32))
        return out

    def sample(self, src, x, y, z) -> int:
        src = np.ascontiguousarray(src, np.uint8)
        return int(self._sample(src, _dims(src), x, y, z))

    def corr_cost(self, frm, to, T) -> float:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        return float(self._corr(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32)))

    def mi_hist(self, frm, to, T):
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        h = np.zeros(256 * 256, np.int32)
        self._mi_hist(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32), h)
        return h.reshape(256, 256)

    def mi_cost(self, frm, to, T) -> float:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        return float(self._mi(frm, _dims(frm), to, _dims(to), np.asarray(T, np.float32)))

    def cost_batch(self, kind: str, frm, to, Ts) -> np.ndarray:
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        out = np.zeros(Ts.shape[0], np.float64)
        self._batch(1 if kind == "mi" else 0, frm, _dims(frm), to, _dims(to), Ts.ravel(),
                    Ts.shape[0], out)
        return out

    def normalize(self, v, upper=255.99):
        out = np.zeros(v.shape, np.uint8)
        if v.dtype == np.int16:
            v = np.ascontiguousarray(v)
            self._norm16(v.ravel(), v.size, _dims(v), out.reshape(-1), upper)
            return out
        v = np.ascontiguousarray(v, np.float32)
        self._norm(v.ravel(), v.size, _dims(v), out.reshape(-1), upper)
        return out

    def downsample(self, v):
        v = np.ascontiguousarray(v, np.uint8)
        s = v.shape
        out = np.zeros(((s[0] + 1) // 2, (s[1] + 1) // 2, (s[2] + 1) // 2), np.uint8)
        od = np.zeros(3, np.int32)
        self._down(v, _dims(v), out, od)
        return out

    def upsample(self, field, out_shape):
        field = _f3(field)
        out = np.zeros(tuple(out_shape) + (3,), np.float32)
        self._up(field, np.array(field.shape[:3][::-1], np.int32), out,
                 np.array(tuple(out_shape)[::-1], np.int32))
        return out

    def compose(self, Is, d):
        Is = np.ascontiguousarray(Is, np.uint8)
        out = np.zeros_like(Is)
        self._compose(Is, _dims(Is), _f3(d), out)
        return out

    def gradient(self, Js, It, kind="r"):
        Js = np.ascontiguousarray(Js, np.uint8)
        It = np.ascontiguousarray(It, np.uint8)
        nd = np.zeros(Js.shape + (3,), np.float32)
        c = self._grad(Js, It, _dims(Js), kind.encode(), nd)
        return nd, float(c)

    def poisson(self, nd):
        nd = _f3(nd).copy()
        self._poisson(nd, np.array(nd.shape[:3][::-1], np.int32))
        return nd

    def max_length(self, d) -> float:
        d = _f3(d)
        return float(self._maxlen(d, np.array(d.shape[:3][::-1], np.int32)))

    def accumulate(self, cur, nd):
        cur = _f3(cur)
        nd = _f3(nd).copy()
        self._acc(cur, nd, np.array(cur.shape[:3][::-1], np.int32))
        return nd

    def smooth(self, d, dd, s):
        d = _f3(d)
        dd = _f3(dd).copy()
        self._smooth(d, dd, np.array(d.shape[:3][::-1], np.int32), s)
        return dd

    def subject_pre(self, img):
        img = np.ascontiguousarray(img, np.float32)
        out = np.zeros(img.shape, np.uint8)
        self._subj(img.reshape(-1), _dims(img), out.reshape(-1))
        return out

    def otsu_median(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._otsu_median(v, v.size))

    def median(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._median(v, v.size))

    def otsu_threshold(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(-1)
        return float(self._otsu(v, v.size))

    def cdm_fixed(self, It, Is, levels=3, level_iterations=48, smoothing=0.8, bandwidth=8):
        It = np.ascontiguousarray(It, np.uint8)
        Is = np.ascontiguousarray(Is, np.uint8)
        field = np.zeros(It.shape + (3,), np.float32)
        self._cdm_fixed(It, Is, _dims(It), levels, level_iterations, smoothing, bandwidth, field)
        return field

    def gaussian(self, img, bandwidth=1):
        out = np.array(img, np.float32, order="C")
        ch = 3 if out.ndim == 4 else 1
        assert out.ndim == 3 or out.shape[3] == 3, "scalar image or vector<3> field"
        self._gauss(out.reshape(-1), np.array(out.shape[:3][::-1], np.int32), ch, int(bandwidth))
        return out

    def improved(self, cost):
        c = np.array(list(cost) + [0.0] * 2, np.float32)
        n = C.c_int(len(cost))
        r = self._improved(c, C.byref(n))
        return bool(r), list(map(float, c[: n.value]))

    def cdm(self, It, Is, resolution=2.0, speed=0.3, smoothing=0.05, min_dimension=16,
            gradient_type="r"):
        if isinstance(It, np.ndarray):
            It, Is = [It], [Is]
        It = [np.ascontiguousarray(v, np.uint8) for v in It]
        Is = [np.ascontiguousarray(v, np.uint8) for v in Is]
        n = len(It)
        pt = (C.c_void_p * n)(*[v.ctypes.data for v in It])
        ps = (C.c_void_p * n)(*[v.ctypes.data for v in Is])
        field = np.zeros(It[0].shape + (3,), np.float32)
        ok = self._cdm(pt, ps, n, _dims(It[0]), resolution, speed, smoothing, min_dimension,
                       gradient_type.encode(), field)
        return bool(ok), field

    def linear(self, frm, fvs, to, tvs, cost="mi", reg_type=15, arg=None):
        frm = np.ascontiguousarray(frm, np.uint8)
        to = np.ascontiguousarray(to, np.uint8)
        a = np.array(arg if arg is not None else [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], np.float32)
        r = self._linear(frm, _dims(frm), np.asarray(fvs, np.float32), to, _dims(to),
                         np.asarray(tvs, np.float32), 1 if cost == "mi" else 0, reg_type, a)
        return float(r), a


def load_oracle() -> Oracle:
    return Oracle()


def load_reference():
    """The compiled reference, or None where it was not built."""
    if not os.path.exists(REF_SO):
        return None
    return Reference()
