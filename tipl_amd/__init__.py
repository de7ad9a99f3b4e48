"""tipl_amd -- MI355X-native registration hot path of TIPL.

The product is libtiplhip.so (HIP kernels for gfx950 behind the C ABI in
include/tiplhip.h) and the header-only C++ mirror of TIPL's entry points in
include/tipl/.  This module is a thin ctypes binding of the C ABI for Python
callers, tests and bench.py.  There is no CPU fallback: without the built
library, or without a GPU, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtiplhip.so")

OK, ERR_INVALID, ERR_DEVICE, ERR_MEMORY, ERR_NO_DEVICE, ERR_TERMINATED, ERR_FAILED = range(7)
U8, I16, F32 = 0, 1, 2
COST_CORR, COST_MI = 0, 1

# symbols declared in include/tiplhip.h
EXPORTS = (
    "tiplhip_last_error", "tiplhip_device_count", "tiplhip_ctx_create", "tiplhip_ctx_destroy",
    "tiplhip_ctx_synchronize", "tiplhip_ctx_set_kernel_timing", "tiplhip_ctx_kernel_time",
    "tiplhip_ctx_set_graphs",
    "tiplhip_timer_start", "tiplhip_timer_stop", "tiplhip_stream_copy", "tiplhip_traffic_probe", "tiplhip_volume_create",
    "tiplhip_volume_create_device", "tiplhip_volume_destroy", "tiplhip_volume_dims",
    "tiplhip_volume_read", "tiplhip_volume_stats", "tiplhip_volume_downsample",
    "tiplhip_resample", "tiplhip_corr_sums", "tiplhip_mi_hist", "tiplhip_cost_batch", "tiplhip_cost_batch_noalloc", "tiplhip_set_launch_shape",
    "tiplhip_cost_batch_params", "tiplhip_affine_matrix", "tiplhip_matrix_inverse",
    "tiplhip_cdm_param_default", "tiplhip_cdm", "tiplhip_cdm_level_iterations",
    "tiplhip_compose_displacement", "tiplhip_cdm_gradient", "tiplhip_cdm_poisson",
    "tiplhip_accumulate_displacement", "tiplhip_cdm_smooth", "tiplhip_upsample_field",
    "tiplhip_gaussian", "tiplhip_cost_batch_pairs", "tiplhip_volume_resample",
    "tiplhip_reg_request_default", "tiplhip_register", "tiplhip_server_create",
    "tiplhip_request_bytes", "tiplhip_server_submit", "tiplhip_server_wait",
    "tiplhip_server_stats_get", "tiplhip_server_destroy", "tiplhip_server_create_ex", "tiplhip_volume_create_subject",
    "tiplhip_autotune", "tiplhip_displacement_to_mapping", "tiplhip_mapping_to_displacement",
    "tiplhip_inv_displacement_to_mapping", "tiplhip_compose_mapping",
    "tiplhip_jacobian_determinant_dis", "tiplhip_jacobian_determinant", "tiplhip_match_fov",
    "tiplhip_correlation", "tiplhip_nifti_read_info", "tiplhip_nifti_read", "tiplhip_nifti_write",
    "tiplhip_cdm2d", "tiplhip_cdm2d_compose", "tiplhip_cdm2d_gradient", "tiplhip_cdm2d_poisson",
    "tiplhip_cdm2d_accumulate", "tiplhip_cdm2d_smooth", "tiplhip_cdm2d_downsample",
    "tiplhip_cdm2d_upsample",
)


class TiplHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tiplhip error {code}: {msg}")
        self.code = code


class PairCosts(C.Structure):
    _fields_ = [("from_", C.c_void_p), ("to", C.c_void_p), ("T", C.c_void_p), ("K", C.c_int),
                ("costs", C.c_void_p)]


class CdmParam(C.Structure):
    """tiplhip_cdm_param (cdm_param, reg/cdm.hpp:36-44, plus the iteration cap)."""
    _fields_ = [("resolution", C.c_float), ("speed", C.c_float), ("smoothing", C.c_float),
                ("min_dimension", C.c_int), ("max_iterations", C.c_int),
                ("gradient_type", C.c_char * 8), ("levels", C.c_int), ("level_iterations", C.c_int),
                ("field_bandwidth", C.c_int)]


class RegRequest(C.Structure):
    """tiplhip_reg_request: one image pair to register (fixed = from / It, moving = to / Is)."""
    _fields_ = [("fixed", C.c_void_p), ("fixed_dtype", C.c_int), ("fixed_dim", C.c_int * 3),
                ("fixed_vs", C.c_float * 3), ("moving", C.c_void_p), ("moving_dtype", C.c_int),
                ("moving_dim", C.c_int * 3), ("moving_vs", C.c_float * 3), ("cost", C.c_int),
                ("bins", C.c_int), ("nonlinear", C.c_int), ("cdm", CdmParam), ("field", C.c_void_p),
                ("preprocess", C.c_int), ("reverse", C.c_int), ("match_fov", C.c_int),
                ("masked_r", C.c_int), ("skip_linear", C.c_int), ("arg", C.c_float * 12),
                ("f2t", C.c_void_p), ("to2from", C.c_void_p), ("from2to", C.c_void_p),
                ("warped", C.c_void_p)]


class RegResult(C.Structure):
    _fields_ = [("status", C.c_int), ("affine", C.c_float * 12), ("T", C.c_float * 12),
                ("cost", C.c_float), ("cdm_iterations", C.c_int), ("evaluations", C.c_int64),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("prep_ms", C.c_double),
                ("affine_ms", C.c_double), ("cdm_ms", C.c_double), ("r_linear", C.c_float),
                ("r_nonlinear", C.c_float)]


class NiftiInfo(C.Structure):
    _fields_ = [("dim", C.c_int * 8), ("pixdim", C.c_float * 8), ("srow", C.c_float * 12),
                ("datatype", C.c_int), ("bitpix", C.c_int), ("scl_slope", C.c_float),
                ("scl_inter", C.c_float), ("sform_code", C.c_int), ("qform_code", C.c_int),
                ("is_mni", C.c_int), ("version", C.c_int), ("big_endian", C.c_int),
                ("flip_swap", C.c_int * 6), ("flip_swap_n", C.c_int), ("descrip", C.c_char * 80)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("running", C.c_int64), ("queued", C.c_int64), ("batch_launches", C.c_int64),
                ("batched_pairs", C.c_int64), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("latency_p50_ms", C.c_double), ("latency_p95_ms", C.c_double),
                ("latency_p99_ms", C.c_double), ("wall_s", C.c_double), ("regs_per_s", C.c_double),
                ("cdm_steps", C.c_int64), ("cdm_fields", C.c_int64)]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libtiplhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m tipl_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(path)
    vp, ip = C.c_void_p, C.c_int
    P = C.POINTER
    sig = {
        "tiplhip_last_error": (C.c_char_p, []),
        "tiplhip_device_count": (ip, [P(ip)]),
        "tiplhip_ctx_create": (ip, [ip, P(vp)]),
        "tiplhip_ctx_destroy": (ip, [vp]),
        "tiplhip_ctx_synchronize": (ip, [vp]),
        "tiplhip_ctx_set_kernel_timing": (ip, [vp, ip]),
        "tiplhip_ctx_set_graphs": (ip, [vp, ip]),
        "tiplhip_ctx_kernel_time": (ip, [vp, P(C.c_double), P(C.c_int64)]),
        "tiplhip_timer_start": (ip, [vp]),
        "tiplhip_timer_stop": (ip, [vp, P(C.c_float)]),
        "tiplhip_stream_copy": (ip, [vp, C.c_size_t, ip, P(C.c_double)]),
        "tiplhip_traffic_probe": (ip, [vp, ip, C.c_size_t, ip, P(C.c_double)]),
        "tiplhip_volume_create": (ip, [vp, vp, ip, P(ip), C.c_float, P(vp)]),
        "tiplhip_volume_create_device": (ip, [vp, vp, ip, P(ip), C.c_float, P(vp)]),
        "tiplhip_volume_destroy": (ip, [vp]),
        "tiplhip_volume_dims": (ip, [vp, P(ip)]),
        "tiplhip_volume_read": (ip, [vp, vp, vp]),
        "tiplhip_volume_stats": (ip, [vp, P(C.c_int64), P(C.c_int64), vp]),
        "tiplhip_volume_downsample": (ip, [vp, vp, P(vp)]),
        "tiplhip_resample": (ip, [vp, vp, P(ip), vp, ip, vp]),
        "tiplhip_corr_sums": (ip, [vp, vp, vp, vp, ip, vp]),
        "tiplhip_mi_hist": (ip, [vp, vp, vp, vp, ip, ip, vp]),
        "tiplhip_cost_batch": (ip, [vp, vp, vp, vp, ip, ip, ip, vp]),
        "tiplhip_cost_batch_noalloc": (ip, [vp, vp, vp, vp, ip, ip, ip, vp]),
        "tiplhip_set_launch_shape": (ip, [vp, vp, vp, ip, ip, ip, ip, ip]),
        "tiplhip_cost_batch_params": (ip, [vp, vp, vp, vp, vp, vp, ip, ip, ip, vp]),
        "tiplhip_affine_matrix": (ip, [vp, P(ip), vp, P(ip), vp, vp]),
        "tiplhip_matrix_inverse": (ip, [vp]),
        "tiplhip_cdm_param_default": (None, [P(CdmParam)]),
        "tiplhip_cdm": (ip, [vp, vp, vp, ip, P(CdmParam), vp, vp, P(ip)]),
        "tiplhip_cdm_level_iterations": (ip, [vp, vp, vp, P(CdmParam), ip, vp]),
        "tiplhip_compose_displacement": (ip, [vp, vp, P(ip), vp, vp]),
        "tiplhip_cdm_gradient": (ip, [vp, vp, vp, P(ip), C.c_char, vp, P(C.c_float)]),
        "tiplhip_cdm_poisson": (ip, [vp, vp, P(ip), ip, P(C.c_float)]),
        "tiplhip_accumulate_displacement": (ip, [vp, vp, vp, P(ip), C.c_float]),
        "tiplhip_cdm_smooth": (ip, [vp, vp, vp, P(ip), C.c_float]),
        "tiplhip_upsample_field": (ip, [vp, vp, P(ip), vp, P(ip)]),
        "tiplhip_cdm2d": (ip, [vp, vp, vp, ip, P(ip), P(CdmParam), vp, vp, P(ip)]),
        "tiplhip_cdm2d_compose": (ip, [vp, vp, P(ip), vp, vp]),
        "tiplhip_cdm2d_gradient": (ip, [vp, vp, vp, P(ip), C.c_char, vp, P(C.c_float)]),
        "tiplhip_cdm2d_poisson": (ip, [vp, vp, P(ip), ip, P(C.c_float)]),
        "tiplhip_cdm2d_accumulate": (ip, [vp, vp, vp, P(ip), C.c_float]),
        "tiplhip_cdm2d_smooth": (ip, [vp, vp, vp, P(ip), C.c_float]),
        "tiplhip_cdm2d_downsample": (ip, [vp, vp, P(ip), vp]),
        "tiplhip_cdm2d_upsample": (ip, [vp, vp, P(ip), vp, P(ip)]),
        "tiplhip_gaussian": (ip, [vp, vp, P(ip), ip, C.c_uint]),
        "tiplhip_cost_batch_pairs": (ip, [vp, P(PairCosts), ip, ip, ip]),
        "tiplhip_volume_resample": (ip, [vp, vp, P(ip), vp, ip, P(vp)]),
        "tiplhip_reg_request_default": (None, [P(RegRequest)]),
        "tiplhip_register": (ip, [vp, P(RegRequest), P(RegResult)]),
        "tiplhip_server_create": (ip, [ip, ip, C.c_size_t, P(vp)]),
        "tiplhip_server_create_ex": (ip, [ip, ip, ip, C.c_size_t, P(vp)]),
        "tiplhip_request_bytes": (C.c_size_t, [P(RegRequest)]),
        "tiplhip_server_submit": (ip, [vp, P(RegRequest), P(C.c_uint64)]),
        "tiplhip_server_wait": (ip, [vp, C.c_uint64, P(RegResult)]),
        "tiplhip_server_stats_get": (ip, [vp, P(ServerStats)]),
        "tiplhip_server_destroy": (ip, [vp]),
        "tiplhip_volume_create_subject": (ip, [vp, vp, ip, P(ip), P(vp)]),
        "tiplhip_autotune": (ip, [vp, vp, vp, vp, ip, ip, ip, ip, P(ip), P(ip), P(C.c_double)]),
        "tiplhip_displacement_to_mapping": (ip, [vp, vp, P(ip), vp, vp]),
        "tiplhip_mapping_to_displacement": (ip, [vp, vp, P(ip), vp]),
        "tiplhip_inv_displacement_to_mapping": (ip, [vp, vp, P(ip), vp, P(ip), vp]),
        "tiplhip_compose_mapping": (ip, [vp, vp, ip, P(ip), vp, P(ip), vp]),
        "tiplhip_jacobian_determinant_dis": (ip, [vp, vp, P(ip), vp]),
        "tiplhip_jacobian_determinant": (ip, [vp, vp, P(ip), vp]),
        "tiplhip_match_fov": (ip, [vp, vp, vp, P(ip)]),
        "tiplhip_correlation": (ip, [vp, vp, vp, ip, P(C.c_double)]),
        "tiplhip_nifti_read_info": (ip, [C.c_char_p, P(NiftiInfo)]),
        "tiplhip_nifti_read": (ip, [C.c_char_p, ip, ip, vp, P(NiftiInfo)]),
        "tiplhip_nifti_write": (ip, [C.c_char_p, vp, ip, P(ip), vp, vp, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != OK:
        msg = load_library().tiplhip_last_error().decode(errors="replace")
        raise TiplHipError(rc, msg)


def _dims(shape) -> "C.Array":
    """x-fastest {w,h,d} from a C-ordered (z, y, x) numpy shape."""
    z, y, x = shape
    return (C.c_int * 3)(x, y, z)


def _dims2(shape) -> "C.Array":
    """x-fastest {w,h} from a C-ordered (y, x) numpy shape."""
    y, x = shape
    return (C.c_int * 2)(x, y)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    n = C.c_int(0)
    _check(load_library().tiplhip_device_count(C.byref(n)))
    return n.value


def affine_matrix(param, from_shape, from_vs, to_shape, to_vs) -> np.ndarray:
    """transformation_matrix(affine_param, from, from_vs, to, to_vs) as 12 floats."""
    lib = load_library()
    p = np.ascontiguousarray(param, np.float32)
    fv = np.ascontiguousarray(from_vs, np.float32)
    tv = np.ascontiguousarray(to_vs, np.float32)
    T = np.zeros(12, np.float32)
    _check(lib.tiplhip_affine_matrix(_ptr(p), _dims(from_shape), _ptr(fv), _dims(to_shape),
                                     _ptr(tv), _ptr(T)))
    return T


def matrix_inverse(T) -> np.ndarray:
    T = np.array(T, np.float32)
    _check(load_library().tiplhip_matrix_inverse(_ptr(T)))
    return T


class Volume:
    """A device-resident uint8 volume (quantised on upload when not uint8)."""

    def __init__(self, ctx: "Context", handle: int, shape):
        self.ctx = ctx
        self.handle = handle
        self.shape = tuple(shape)

    def read(self) -> np.ndarray:
        out = np.empty(self.shape, np.uint8)
        _check(load_library().tiplhip_volume_read(self.ctx.handle, self.handle, _ptr(out)))
        return out

    def stats(self):
        s, s2 = C.c_int64(0), C.c_int64(0)
        h = np.zeros(256, np.uint32)
        _check(load_library().tiplhip_volume_stats(self.handle, C.byref(s), C.byref(s2), _ptr(h)))
        return s.value, s2.value, h

    def downsample(self) -> "Volume":
        out = C.c_void_p()
        _check(load_library().tiplhip_volume_downsample(self.ctx.handle, self.handle, C.byref(out)))
        z, y, x = self.shape
        return Volume(self.ctx, out.value, ((z + 1) // 2, (y + 1) // 2, (x + 1) // 2))

    def free(self):
        if self.handle:
            _check(load_library().tiplhip_volume_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                load_library().tiplhip_volume_destroy(self.handle)
        except Exception:
            pass


class Context:
    """A device context: one GPU, one HIP stream, reusable scratch memory."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.tiplhip_ctx_create(device, C.byref(h)))
        self.handle = h.value
        self.device = device

    def close(self):
        if self.handle:
            _check(load_library().tiplhip_ctx_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        _check(load_library().tiplhip_ctx_synchronize(self.handle))

    def set_graphs(self, enable: bool):
        """Replays captured HIP graphs for cost launches of a shape seen before."""
        _check(load_library().tiplhip_ctx_set_graphs(self.handle, int(enable)))

    # -- timing ------------------------------------------------------------
    def set_kernel_timing(self, enable: bool):
        _check(load_library().tiplhip_ctx_set_kernel_timing(self.handle, int(enable)))

    def kernel_time(self):
        ms, n = C.c_double(0), C.c_int64(0)
        _check(load_library().tiplhip_ctx_kernel_time(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timer_start(self):
        _check(load_library().tiplhip_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_float(0)
        _check(load_library().tiplhip_timer_stop(self.handle, C.byref(ms)))
        return ms.value

    def stream_copy(self, nbytes: int, reps: int = 20) -> float:
        """Achievable HBM bandwidth of a plain device copy, GB/s (read + write)."""
        g = C.c_double(0)
        _check(load_library().tiplhip_stream_copy(self.handle, nbytes, reps, C.byref(g)))
        return g.value

    # -- volumes -----------------------------------------------------------
    def traffic_probe(self, kind: int, nbytes: int, reps: int = 20) -> float:
        """GB/s of a kernel of known traffic: 0 copy, 1 16-byte reads, 2 float3 reads."""
        g = C.c_double(0)
        _check(load_library().tiplhip_traffic_probe(self.handle, kind, nbytes, reps, C.byref(g)))
        return g.value

    def volume(self, data: np.ndarray, upper_limit: float = 255.99) -> Volume:
        if data.dtype == np.uint8:
            dt = U8
        elif data.dtype == np.int16:
            dt = I16
        else:
            data = data.astype(np.float32, copy=False)
            dt = F32
        data = np.ascontiguousarray(data)
        h = C.c_void_p()
        _check(load_library().tiplhip_volume_create(self.handle, _ptr(data), dt, _dims(data.shape),
                                                    C.c_float(upper_limit), C.byref(h)))
        return Volume(self, h.value, data.shape)

    def volume_subject(self, data: np.ndarray) -> Volume:
        """reg::subject_image_pre on the device (gaussian, Otsu-median normalisation, uint8)."""
        dt = {np.dtype(np.uint8): U8, np.dtype(np.int16): I16}.get(data.dtype, F32)
        data = np.ascontiguousarray(data if dt != F32 else data.astype(np.float32, copy=False))
        h = C.c_void_p()
        _check(load_library().tiplhip_volume_create_subject(self.handle, _ptr(data), dt,
                                                            _dims(data.shape), C.byref(h)))
        return Volume(self, h.value, data.shape)

    # -- affine --------------------------------------------------------------
    def resample(self, src: Volume, out_shape, T, crop: bool = True) -> np.ndarray:
        out = np.empty(tuple(out_shape), np.uint8)
        Tm = np.ascontiguousarray(T, np.float32)
        _check(load_library().tiplhip_resample(self.handle, src.handle, _dims(out_shape), _ptr(Tm),
                                               int(crop), _ptr(out)))
        return out

    def corr_sums(self, frm: Volume, to: Volume, Ts) -> np.ndarray:
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        out = np.zeros((Ts.shape[0], 5), np.int64)
        _check(load_library().tiplhip_corr_sums(self.handle, frm.handle, to.handle, _ptr(Ts),
                                                Ts.shape[0], _ptr(out)))
        return out

    def mi_hist(self, frm: Volume, to: Volume, Ts, bins: int = 256) -> np.ndarray:
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        out = np.zeros((Ts.shape[0], bins, bins), np.int32)
        _check(load_library().tiplhip_mi_hist(self.handle, frm.handle, to.handle, _ptr(Ts),
                                              Ts.shape[0], bins, _ptr(out)))
        return out

    def cost_batch(self, frm: Volume, to: Volume, Ts, cost: str = "mi", bins: int = 256) -> np.ndarray:
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        out = np.zeros(Ts.shape[0], np.float64)
        _check(load_library().tiplhip_cost_batch(self.handle, frm.handle, to.handle, _ptr(Ts),
                                                 Ts.shape[0], COST_MI if cost == "mi" else COST_CORR,
                                                 bins, _ptr(out)))
        return out

    def cost_batch_noalloc(self, frm: Volume, to: Volume, Ts, cost: str = "mi",
                           bins: int = 256) -> np.ndarray:
        """cost_batch through the entry point that allocates no device memory
        (K <= 256; 256 bins: K <= 79; the sampled volume packed by an earlier
        cost_batch)."""
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        out = np.zeros(Ts.shape[0], np.float64)
        _check(load_library().tiplhip_cost_batch_noalloc(
            self.handle, frm.handle, to.handle, _ptr(Ts), Ts.shape[0],
            COST_MI if cost == "mi" else COST_CORR, bins, _ptr(out)))
        return out

    def set_launch_shape(self, frm: Volume, to: Volume, K: int, cost: str = "mi", bins: int = 256,
                         zc: int = 16, per_cu: int = 2):
        """Pins the launch shape of later single-pair launches with this key."""
        _check(load_library().tiplhip_set_launch_shape(
            self.handle, frm.handle, to.handle, K, COST_MI if cost == "mi" else COST_CORR, bins,
            zc, per_cu))

    def autotune(self, frm: Volume, to: Volume, Ts, cost: str = "mi", bins: int = 256, reps: int = 3):
        """Picks the cost kernel's launch shape for this pair, K and cost (kept on
        the context); returns (planes per item, workgroups per CU, best kernel ms)."""
        Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
        zc, pc, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(load_library().tiplhip_autotune(self.handle, frm.handle, to.handle, _ptr(Ts),
                                               Ts.shape[0], COST_MI if cost == "mi" else COST_CORR,
                                               bins, reps, C.byref(zc), C.byref(pc), C.byref(ms)))
        return zc.value, pc.value, ms.value

    def cost_batch_pairs(self, pairs, cost: str = "mi", bins: int = 256) -> list:
        """One launch over several image pairs: pairs = [(from, to, Ts (K x 12)), ...];
        returns one cost array per pair."""
        keep, outs = [], []
        arr = (PairCosts * len(pairs))()
        for i, (frm, to, Ts) in enumerate(pairs):
            Ts = np.ascontiguousarray(Ts, np.float32).reshape(-1, 12)
            out = np.zeros(Ts.shape[0], np.float64)
            keep.append(Ts)
            outs.append(out)
            arr[i] = PairCosts(frm.handle, to.handle, Ts.ctypes.data, Ts.shape[0], out.ctypes.data)
        _check(load_library().tiplhip_cost_batch_pairs(self.handle, arr, len(pairs),
                                                       COST_MI if cost == "mi" else COST_CORR, bins))
        return outs

    def cost_batch_params(self, frm: Volume, from_vs, to: Volume, to_vs, params, cost="mi",
                          bins=256) -> np.ndarray:
        P = np.ascontiguousarray(params, np.float32).reshape(-1, 12)
        fv = np.ascontiguousarray(from_vs, np.float32)
        tv = np.ascontiguousarray(to_vs, np.float32)
        out = np.zeros(P.shape[0], np.float64)
        _check(load_library().tiplhip_cost_batch_params(
            self.handle, frm.handle, _ptr(fv), to.handle, _ptr(tv), _ptr(P), P.shape[0],
            COST_MI if cost == "mi" else COST_CORR, bins, _ptr(out)))
        return out

    # -- CDM -----------------------------------------------------------------
    @staticmethod
    def cdm_param(**kw) -> CdmParam:
        p = CdmParam()
        load_library().tiplhip_cdm_param_default(C.byref(p))
        gt = kw.pop("gradient_type", None)
        for k, v in kw.items():
            setattr(p, k, v)
        if gt is not None:
            p.gradient_type = (gt * 8)[:8].encode() if len(gt) == 1 else gt.encode()[:8].ljust(8, b"r")
        return p

    def cdm(self, It: Sequence[Volume], Is: Sequence[Volume], param: Optional[CdmParam] = None,
            terminated: Optional[C.c_uint8] = None):
        if isinstance(It, Volume):
            It, Is = [It], [Is]
        n = len(It)
        at = (C.c_void_p * n)(*[v.handle for v in It])
        as_ = (C.c_void_p * n)(*[v.handle for v in Is])
        field = np.zeros(It[0].shape + (3,), np.float32)
        iters = C.c_int(0)
        p = param if param is not None else self.cdm_param()
        rc = load_library().tiplhip_cdm(self.handle, at, as_, n, C.byref(p), _ptr(field),
                                        C.byref(terminated) if terminated is not None else None,
                                        C.byref(iters))
        return rc, field, iters.value

    def cdm_level_iterations(self, It: Volume, Is: Volume, iterations: int,
                             param: Optional[CdmParam] = None, want_field: bool = False):
        p = param if param is not None else self.cdm_param()
        field = np.zeros(It.shape + (3,), np.float32) if want_field else None
        _check(load_library().tiplhip_cdm_level_iterations(
            self.handle, It.handle, Is.handle, C.byref(p), iterations,
            _ptr(field) if want_field else None))
        return field

    # -- post-registration warping (numerical/dif.hpp) -----------------------
    def displacement_to_mapping(self, dis: np.ndarray, T=None) -> np.ndarray:
        dis = np.ascontiguousarray(dis, np.float32)
        out = np.empty_like(dis)
        Tm = None if T is None else np.ascontiguousarray(T, np.float32)
        _check(load_library().tiplhip_displacement_to_mapping(
            self.handle, _ptr(dis), _dims(dis.shape[:3]), None if Tm is None else _ptr(Tm), _ptr(out)))
        return out

    def mapping_to_displacement(self, mapping: np.ndarray) -> np.ndarray:
        mapping = np.ascontiguousarray(mapping, np.float32)
        out = np.empty_like(mapping)
        _check(load_library().tiplhip_mapping_to_displacement(self.handle, _ptr(mapping),
                                                              _dims(mapping.shape[:3]), _ptr(out)))
        return out

    def inv_displacement_to_mapping(self, inv_dis: np.ndarray, T, out_shape) -> np.ndarray:
        inv_dis = np.ascontiguousarray(inv_dis, np.float32)
        Tm = np.ascontiguousarray(T, np.float32)
        out = np.empty(tuple(out_shape) + (3,), np.float32)
        _check(load_library().tiplhip_inv_displacement_to_mapping(
            self.handle, _ptr(inv_dis), _dims(inv_dis.shape[:3]), _ptr(Tm), _dims(out_shape), _ptr(out)))
        return out

    def compose_mapping(self, src: np.ndarray, mapping: np.ndarray) -> np.ndarray:
        mapping = np.ascontiguousarray(mapping, np.float32)
        if src.dtype == np.uint8:
            src, dt = np.ascontiguousarray(src), U8
        else:
            src, dt = np.ascontiguousarray(src, np.float32), F32
        out = np.empty(mapping.shape[:3], src.dtype)
        _check(load_library().tiplhip_compose_mapping(self.handle, _ptr(src), dt, _dims(src.shape),
                                                      _ptr(mapping), _dims(mapping.shape[:3]),
                                                      _ptr(out)))
        return out

    def jacobian_determinant_dis(self, dis: np.ndarray) -> np.ndarray:
        dis = np.ascontiguousarray(dis, np.float32)
        out = np.empty(dis.shape[:3], np.float32)
        _check(load_library().tiplhip_jacobian_determinant_dis(self.handle, _ptr(dis),
                                                               _dims(dis.shape[:3]), _ptr(out)))
        return out

    def jacobian_determinant(self, mapping: np.ndarray) -> np.ndarray:
        mapping = np.ascontiguousarray(mapping, np.float32)
        out = np.empty(mapping.shape[:3], np.float32)
        _check(load_library().tiplhip_jacobian_determinant(self.handle, _ptr(mapping),
                                                           _dims(mapping.shape[:3]), _ptr(out)))
        return out

    def match_fov(self, It: Volume, T, moving_shape):
        """mm_reg::It_match_fov in place on a device volume."""
        Tm = np.ascontiguousarray(T, np.float32)
        _check(load_library().tiplhip_match_fov(self.handle, It.handle, _ptr(Tm), _dims(moving_shape)))

    def correlation(self, x: Volume, y: Volume, masked: bool = True) -> float:
        r = C.c_double(0)
        _check(load_library().tiplhip_correlation(self.handle, x.handle, y.handle, int(bool(masked)),
                                                  C.byref(r)))
        return r.value

    def compose(self, Is: np.ndarray, d: np.ndarray) -> np.ndarray:
        Is = np.ascontiguousarray(Is, np.uint8)
        d = np.ascontiguousarray(d, np.float32)
        out = np.empty_like(Is)
        _check(load_library().tiplhip_compose_displacement(self.handle, _ptr(Is), _dims(Is.shape),
                                                           _ptr(d), _ptr(out)))
        return out

    def gradient(self, Js: np.ndarray, It: np.ndarray, kind: str = "r"):
        Js = np.ascontiguousarray(Js, np.uint8)
        It = np.ascontiguousarray(It, np.uint8)
        nd = np.zeros(Js.shape + (3,), np.float32)
        c = C.c_float(0)
        _check(load_library().tiplhip_cdm_gradient(self.handle, _ptr(Js), _ptr(It), _dims(Js.shape),
                                                   kind.encode(), _ptr(nd), C.byref(c)))
        return nd, c.value

    def poisson(self, nd: np.ndarray, sweeps: int = 12):
        nd = np.array(nd, np.float32, order="C")
        ml = C.c_float(0)
        _check(load_library().tiplhip_cdm_poisson(self.handle, _ptr(nd), _dims(nd.shape[:3]), sweeps,
                                                  C.byref(ml)))
        return nd, ml.value

    def accumulate(self, cur: np.ndarray, nd: np.ndarray, scale: float = 1.0) -> np.ndarray:
        cur = np.ascontiguousarray(cur, np.float32)
        nd = np.array(nd, np.float32, order="C")
        _check(load_library().tiplhip_accumulate_displacement(self.handle, _ptr(cur), _ptr(nd),
                                                              _dims(cur.shape[:3]), C.c_float(scale)))
        return nd

    def smooth(self, d: np.ndarray, dd: np.ndarray, smoothing: float) -> np.ndarray:
        d = np.ascontiguousarray(d, np.float32)
        dd = np.array(dd, np.float32, order="C")
        _check(load_library().tiplhip_cdm_smooth(self.handle, _ptr(d), _ptr(dd), _dims(d.shape[:3]),
                                                 C.c_float(smoothing)))
        return dd

    def upsample(self, field: np.ndarray, out_shape) -> np.ndarray:
        field = np.ascontiguousarray(field, np.float32)
        out = np.zeros(tuple(out_shape) + (3,), np.float32)
        _check(load_library().tiplhip_upsample_field(self.handle, _ptr(field), _dims(field.shape[:3]),
                                                     _ptr(out), _dims(out_shape)))
        return out

    # -- 2-D CDM (cdm<out,uint8,2>): (h, w) uint8 images, (h, w, 2) fields --
    def cdm2d(self, It, Is, param: Optional[CdmParam] = None,
              terminated: Optional[C.c_uint8] = None):
        """tipl::reg::cdm on 2-D uint8 host images (one or a list of
        modalities); returns (status, field, iterations)."""
        if isinstance(It, np.ndarray):
            It, Is = [It], [Is]
        It = [np.ascontiguousarray(v, np.uint8) for v in It]
        Is = [np.ascontiguousarray(v, np.uint8) for v in Is]
        n = len(It)
        at = (C.c_void_p * n)(*[v.ctypes.data for v in It])
        as_ = (C.c_void_p * n)(*[v.ctypes.data for v in Is])
        field = np.zeros(It[0].shape + (2,), np.float32)
        iters = C.c_int(0)
        p = param if param is not None else self.cdm_param()
        rc = load_library().tiplhip_cdm2d(self.handle, at, as_, n, _dims2(It[0].shape), C.byref(p),
                                          _ptr(field),
                                          C.byref(terminated) if terminated is not None else None,
                                          C.byref(iters))
        return rc, field, iters.value

    def compose2d(self, Is: np.ndarray, d: np.ndarray) -> np.ndarray:
        Is = np.ascontiguousarray(Is, np.uint8)
        d = np.ascontiguousarray(d, np.float32)
        out = np.empty_like(Is)
        _check(load_library().tiplhip_cdm2d_compose(self.handle, _ptr(Is), _dims2(Is.shape), _ptr(d),
                                                     _ptr(out)))
        return out

    def gradient2d(self, Js: np.ndarray, It: np.ndarray, kind: str = "r"):
        Js = np.ascontiguousarray(Js, np.uint8)
        It = np.ascontiguousarray(It, np.uint8)
        nd = np.zeros(Js.shape + (2,), np.float32)
        c = C.c_float(0)
        _check(load_library().tiplhip_cdm2d_gradient(self.handle, _ptr(Js), _ptr(It), _dims2(Js.shape),
                                                      kind.encode(), _ptr(nd), C.byref(c)))
        return nd, c.value

    def poisson2d(self, nd: np.ndarray, sweeps: int = 12):
        nd = np.array(nd, np.float32, order="C")
        ml = C.c_float(0)
        _check(load_library().tiplhip_cdm2d_poisson(self.handle, _ptr(nd), _dims2(nd.shape[:2]), sweeps,
                                                     C.byref(ml)))
        return nd, ml.value

    def accumulate2d(self, cur: np.ndarray, nd: np.ndarray, scale: float = 1.0) -> np.ndarray:
        cur = np.ascontiguousarray(cur, np.float32)
        nd = np.array(nd, np.float32, order="C")
        _check(load_library().tiplhip_cdm2d_accumulate(self.handle, _ptr(cur), _ptr(nd),
                                                        _dims2(cur.shape[:2]), C.c_float(scale)))
        return nd

    def smooth2d(self, d: np.ndarray, dd: np.ndarray, smoothing: float) -> np.ndarray:
        d = np.ascontiguousarray(d, np.float32)
        dd = np.array(dd, np.float32, order="C")
        _check(load_library().tiplhip_cdm2d_smooth(self.handle, _ptr(d), _ptr(dd), _dims2(d.shape[:2]),
                                                    C.c_float(smoothing)))
        return dd

    def downsample2d(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img, np.uint8)
        out = np.empty(((img.shape[0] + 1) // 2, (img.shape[1] + 1) // 2), np.uint8)
        _check(load_library().tiplhip_cdm2d_downsample(self.handle, _ptr(img), _dims2(img.shape),
                                                        _ptr(out)))
        return out

    def upsample2d(self, field: np.ndarray, out_shape) -> np.ndarray:
        field = np.ascontiguousarray(field, np.float32)
        out = np.zeros(tuple(out_shape) + (2,), np.float32)
        _check(load_library().tiplhip_cdm2d_upsample(self.handle, _ptr(field), _dims2(field.shape[:2]),
                                                      _ptr(out), _dims2(out_shape)))
        return out

    def gaussian(self, img: np.ndarray, bandwidth: int = 1) -> np.ndarray:
        """filter::gaussian of a float32 (z, y, x) image or (z, y, x, 3) field; returns a copy."""
        out = np.array(img, np.float32, order="C")
        ch = out.shape[3] if out.ndim == 4 else 1
        _check(load_library().tiplhip_gaussian(self.handle, _ptr(out), _dims(out.shape[:3]), ch,
                                               C.c_uint(bandwidth)))
        return out


# ---------------------------------------------------------------------------
# Registration of image pairs: the plain path and the serving loop
# ---------------------------------------------------------------------------
_DTYPES = {np.dtype(np.uint8): U8, np.dtype(np.int16): I16, np.dtype(np.float32): F32}


def make_request(fixed, moving, fixed_vs=(1, 1, 1), moving_vs=(1, 1, 1), cost="mi", bins=256,
                 nonlinear=False, want_field=True, preprocess="linear", reverse=False,
                 match_fov=True, masked_r=True, arg=None, outputs=(), **cdm):
    """A RegRequest over numpy (z, y, x) volumes of uint8 / int16 / float32.
    preprocess: 'linear' quantises both with normalize_upper_lower2 (255.99);
    'mm_reg' uses template_image_pre for fixed and subject_image_pre for moving.
    reverse / match_fov / masked_r: mm_reg's nonlinear options; arg: a given
    affine_param (skips linear()).  outputs: any of 'f2t', 'to2from',
    'from2to', 'warped'.  Returns (request, keep) where keep holds the arrays
    the request points to and the requested outputs."""
    fixed = np.ascontiguousarray(fixed)
    moving = np.ascontiguousarray(moving)
    if fixed.dtype not in _DTYPES or moving.dtype not in _DTYPES:
        raise TypeError("volumes must be uint8, int16 or float32")
    r = RegRequest()
    load_library().tiplhip_reg_request_default(C.byref(r))
    r.fixed, r.fixed_dtype = fixed.ctypes.data, _DTYPES[fixed.dtype]
    r.moving, r.moving_dtype = moving.ctypes.data, _DTYPES[moving.dtype]
    r.fixed_dim[:] = list(fixed.shape[::-1])
    r.moving_dim[:] = list(moving.shape[::-1])
    r.fixed_vs[:] = list(map(float, fixed_vs))
    r.moving_vs[:] = list(map(float, moving_vs))
    r.cost = COST_MI if cost == "mi" else COST_CORR
    r.bins = int(bins)
    r.nonlinear = int(bool(nonlinear))
    if preprocess not in ("linear", "mm_reg"):
        raise ValueError("preprocess is 'linear' or 'mm_reg'")
    r.preprocess = 1 if preprocess == "mm_reg" else 0
    gt = cdm.pop("gradient_type", None)
    for k, v in cdm.items():
        setattr(r.cdm, k, v)
    if gt is not None:
        r.cdm.gradient_type = (gt * 8)[:8].encode() if len(gt) == 1 else gt.encode()[:8].ljust(8, b"r")
    r.reverse = int(bool(reverse))
    r.match_fov = int(bool(match_fov))
    r.masked_r = int(bool(masked_r))
    if arg is not None:
        r.skip_linear = 1
        r.arg[:] = list(map(float, np.asarray(arg, np.float32)))
    keep = {"fixed": fixed, "moving": moving, "field": None}
    if nonlinear and want_field:
        keep["field"] = np.zeros(fixed.shape + (3,), np.float32)
        r.field = keep["field"].ctypes.data
    shapes = {"f2t": fixed.shape + (3,), "to2from": fixed.shape + (3,),
              "from2to": moving.shape + (3,), "warped": fixed.shape}
    for name in outputs:
        a = np.zeros(shapes[name], np.uint8 if name == "warped" else np.float32)
        setattr(r, name, a.ctypes.data)
        keep[name] = a
    return r, keep


def _result(res: RegResult, keep):
    out = {"status": res.status, "affine": np.array(res.affine[:], np.float32),
           "T": np.array(res.T[:], np.float32), "cost": float(res.cost),
           "cdm_iterations": int(res.cdm_iterations), "evaluations": int(res.evaluations),
           "queue_ms": float(res.queue_ms), "run_ms": float(res.run_ms),
           "prep_ms": float(res.prep_ms), "affine_ms": float(res.affine_ms),
           "cdm_ms": float(res.cdm_ms), "r_linear": float(res.r_linear),
           "r_nonlinear": float(res.r_nonlinear)}
    for name in ("field", "f2t", "to2from", "from2to", "warped"):
        out[name] = keep.get(name)
    return out


def register(ctx: "Context", fixed, moving, **kw) -> dict:
    """The plain single-pair path (tiplhip_register) on ctx's thread and stream."""
    req, keep = make_request(fixed, moving, **kw)
    res = RegResult()
    _check(load_library().tiplhip_register(ctx.handle, C.byref(req), C.byref(res)))
    return _result(res, keep)


def request_bytes(fixed, moving, **kw) -> int:
    req, _ = make_request(fixed, moving, want_field=False, **kw)
    return int(load_library().tiplhip_request_bytes(C.byref(req)))


class Server:
    """The serving loop (tiplhip_server_*): slots, HBM-budget admission, one cost batcher."""

    def __init__(self, device: int = 0, slots: int = 8, hbm_budget: int = 64 << 30, launchers: int = -1):
        h = C.c_void_p()
        _check(load_library().tiplhip_server_create_ex(device, slots, launchers, C.c_size_t(hbm_budget),
                                                       C.byref(h)))
        self.handle = h.value
        self._keep = {}

    def submit(self, fixed, moving, **kw) -> int:
        req, keep = make_request(fixed, moving, **kw)
        t = C.c_uint64(0)
        _check(load_library().tiplhip_server_submit(self.handle, C.byref(req), C.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def wait(self, ticket: int) -> dict:
        res = RegResult()
        keep = self._keep.pop(ticket, {"field": None})
        _check(load_library().tiplhip_server_wait(self.handle, C.c_uint64(ticket), C.byref(res)))
        return _result(res, keep)

    def stats(self) -> dict:
        st = ServerStats()
        _check(load_library().tiplhip_server_stats_get(self.handle, C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            _check(load_library().tiplhip_server_destroy(self.handle))
            self.handle = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# -- NIfTI files (host) ---------------------------------------------------------
def _info_dict(info: NiftiInfo) -> dict:
    return {"dim": list(info.dim), "vs": np.array(info.pixdim[1:4], np.float32),
            "srow": np.array(info.srow[:], np.float32), "datatype": info.datatype,
            "bitpix": info.bitpix, "scl_slope": info.scl_slope, "scl_inter": info.scl_inter,
            "sform_code": info.sform_code, "qform_code": info.qform_code,
            "is_mni": bool(info.is_mni), "version": info.version, "big_endian": bool(info.big_endian),
            "flip_swap": list(info.flip_swap[:info.flip_swap_n]),
            "descrip": info.descrip.decode(errors="replace")}


def nifti_info(path) -> dict:
    """Header of a NIfTI / Analyze file after the reference's LPS reorientation."""
    info = NiftiInfo()
    _check(load_library().tiplhip_nifti_read_info(os.fsencode(str(path)), C.byref(info)))
    return _info_dict(info)


def nifti_read(path, volume: int = 0, dtype=np.float32):
    """One volume as a (z, y, x) array in LPS orientation (tipl::io::nifti's
    `>>`), and the header; dtype uint8, int16 or float32."""
    info = NiftiInfo()
    lib = load_library()
    _check(lib.tiplhip_nifti_read_info(os.fsencode(str(path)), C.byref(info)))
    dt = np.dtype(dtype)
    code = {np.dtype(np.uint8): U8, np.dtype(np.int16): I16, np.dtype(np.float32): F32}[dt]
    out = np.empty((info.dim[3], info.dim[2], info.dim[1]), dt)
    _check(lib.tiplhip_nifti_read(os.fsencode(str(path)), int(volume), code, _ptr(out), C.byref(info)))
    return out, _info_dict(info)


def nifti_write(path, data: np.ndarray, vs=(1.0, 1.0, 1.0), srow=None, is_mni: bool = False):
    """Writes a (z, y, x) uint8 / int16 / float32 volume as NIfTI-1 (gzip for .gz)."""
    data = np.ascontiguousarray(data)
    code = {np.dtype(np.uint8): U8, np.dtype(np.int16): I16, np.dtype(np.float32): F32}[data.dtype]
    if srow is None:
        srow = np.array([vs[0], 0, 0, 0, 0, vs[1], 0, 0, 0, 0, vs[2], 0], np.float32)
    v = np.ascontiguousarray(vs, np.float32)
    r = np.ascontiguousarray(srow, np.float32)
    _check(load_library().tiplhip_nifti_write(os.fsencode(str(path)), _ptr(data), code,
                                              _dims(data.shape), _ptr(v), _ptr(r), int(bool(is_mni))))


def register_files(ctx: "Context", template_path, subject_path, volume: int = 0, **kw) -> dict:
    """mm_reg's file entry (reg/mm_reg.hpp:100-152 load_reg_image, then the
    registration): the template (fixed) and the subject (moving) read from
    NIfTI files in LPS orientation with their voxel sizes, preprocessed as
    mm_reg does (template_image_pre / subject_image_pre, on the device), then
    tiplhip_register.  Extra keywords go to make_request."""
    fixed, fi = nifti_read(template_path, volume)
    moving, mi = nifti_read(subject_path, volume)
    kw.setdefault("preprocess", "mm_reg")
    res = register(ctx, fixed, moving, fixed_vs=tuple(fi["vs"]), moving_vs=tuple(mi["vs"]), **kw)
    res["template_info"], res["subject_info"] = fi, mi
    return res
