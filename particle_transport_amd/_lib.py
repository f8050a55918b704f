"""ctypes binding of libpt_hip.so (the C ABI declared in include/pt_hip.h).

The library must exist (build it with ``python -m particle_transport_amd.build``):
there is no CPU fallback, and every call raises ``PTError`` with the
library's own message when a status is non-zero.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpt_hip.so")

PT_OK, PT_ERR_INVALID, PT_ERR_HIP, PT_ERR_MEMORY, PT_ERR_STATE = 0, 1, 2, 3, 4
PT_F64, PT_F32 = 0, 1
STORAGE = {"f64_planar": 0, "f32_planar": 1, "f32_brick": 2, "bf16_brick": 3, "f16_brick": 4}
PUSHER = {"ref_f64": 0, "plain_f32": 1, "fast": 2}
OUT = {"none": 0, "phase": 1, "phase-space": 1, "mu": 2}

# Every symbol include/pt_hip.h declares.
EXPORTS = (
    "pt_abi_version", "pt_last_error", "pt_ctx_create", "pt_ctx_destroy", "pt_ctx_bytes_in_use", "pt_device_memory",
    "pt_field_create", "pt_field_destroy", "pt_field_cell_table", "pt_field_mean_abs_b", "pt_mean_abs_b",
    "pt_ens_create", "pt_ens_create_injected", "pt_inject", "pt_ens_destroy", "pt_ens_advance",
    "pt_ens_get_state", "pt_ens_last_timing", "pt_ens_sort",
    "pt_mu_from_phase", "pt_mu_from_phase_layout", "pt_dmumu_accumulate", "pt_dmumu_accumulate_noalloc", "pt_dpp_accumulate",
    "pt_dpp_accumulate_pmag",
    "pt_dpp_accumulate_dprel",
    "pt_energy_snapshots", "pt_buffer_alloc", "pt_buffer_free", "pt_memcpy", "pt_memcpy_2d", "pt_host_alloc", "pt_host_free",
    "pt_serve_create", "pt_serve_destroy", "pt_serve_add_field", "pt_serve_submit", "pt_serve_step",
    "pt_serve_drain", "pt_serve_take_done", "pt_serve_poll", "pt_serve_fetch", "pt_serve_copy_cache", "pt_serve_release",
    "pt_serve_stats",
)


class PTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libpt_hip error {code}: {msg}")
        self.code = code


class AdvanceParams(ctypes.Structure):
    _fields_ = [("dt_s", ctypes.c_double), ("charge_C", ctypes.c_double), ("mass_kg", ctypes.c_double),
                ("first_step", ctypes.c_int64), ("n_steps", ctypes.c_int64), ("total_steps", ctypes.c_int64),
                ("save_stride", ctypes.c_int64), ("out_kind", ctypes.c_int32),
                ("steps_per_launch", ctypes.c_int32), ("sort_every_steps", ctypes.c_int64)]


class InjectParams(ctypes.Structure):
    """pt_inject_params: injection settings of init_particles."""
    _fields_ = [("seed", ctypes.c_uint64), ("fixed_position", ctypes.c_int32), ("fixed_mu", ctypes.c_int32),
                ("position_m", ctypes.c_double * 3), ("mu0", ctypes.c_double), ("p_magnitude", ctypes.c_double)]


_lib = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_int = ctypes.c_int


def load():
    """Load libpt_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m particle_transport_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "pt_abi_version": (_int, []),
        "pt_last_error": (ctypes.c_char_p, []),
        "pt_ctx_create": (_int, [_int, ctypes.POINTER(_vp)]),
        "pt_ctx_destroy": (_int, [_vp]),
        "pt_ctx_bytes_in_use": (_int, [_vp, _i64p]),
        "pt_device_memory": (_int, [_vp, _i64p, _i64p]),
        "pt_field_create": (_int, [_vp, _i64, _i64, _i64, ctypes.c_double, _int, _int, _int,
                                   ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
        "pt_field_destroy": (_int, [_vp]),
        "pt_field_cell_table": (_int, [_vp, _int]),
        "pt_field_mean_abs_b": (_int, [_vp, _dp]),
        "pt_mean_abs_b": (_int, [_vp, _i64, _int, _int, _vp, _vp, _vp, _dp]),
        "pt_ens_create": (_int, [_vp, _vp, _int, _i64, _int, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 _int, ctypes.POINTER(_vp)]),
        "pt_ens_create_injected": (_int, [_vp, _vp, _int, _i64, ctypes.POINTER(InjectParams), _int,
                                          ctypes.POINTER(_vp)]),
        "pt_inject": (_int, [_vp, _vp, _i64, _i64, ctypes.POINTER(InjectParams), _int, _dp]),
        "pt_ens_destroy": (_int, [_vp]),
        "pt_ens_advance": (_int, [_vp, _vp, ctypes.POINTER(AdvanceParams), _vp, _vp, _int, _int, _i64,
                                  _i64p, _i64p]),
        "pt_ens_get_state": (_int, [_vp, ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(ctypes.c_uint8)]),
        "pt_ens_last_timing": (_int, [_vp, _dp, _i64p]),
        "pt_ens_sort": (_int, [_vp]),
        "pt_mu_from_phase": (_int, [_vp, _vp, _i64, _i64, _int, _int, _vp, _vp, _int, _int, _vp]),
        "pt_mu_from_phase_layout": (_int, [_vp, _vp, _i64, _i64, _int, _int, _vp, _vp, _int, _int, _vp, _int]),
        "pt_dmumu_accumulate": (_int, [_vp, _fp, _int, _i64, _i64, _i64, _i64, _i64p, _i64p, _i64, _dp, _i64,
                                       _int, _int, _i64p, _dp, _dp, _dp, _i64p, _i64p, _dp, _dp]),
        "pt_dmumu_accumulate_noalloc": (_int, [_vp, _fp, _i64, _i64, _i64, _i64, _i64p, _i64p, _i64, _dp, _i64,
                                               _int, _int, _i64p, _dp, _dp, _dp, _i64p, _i64p, _dp, _dp]),
        "pt_dpp_accumulate": (_int, [_vp, _fp, _int, _i64, _i64, _i64, _i64p, _i64p, _i64, ctypes.c_double,
                                     _i64p, _dp, _dp, _dp, _dp]),
        "pt_energy_snapshots": (_int, [_vp, _fp, _int, _i64, _i64, _i64, _i64p, _i64p, _i64, _dp, _i64]),
        "pt_buffer_alloc": (_int, [_vp, _i64, ctypes.POINTER(_vp)]),
        "pt_buffer_free": (_int, [_vp, _vp, _i64]),
        "pt_memcpy": (_int, [_vp, _vp, _vp, _i64, _int]),
        "pt_memcpy_2d": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _int]),
        "pt_host_alloc": (_int, [_i64, ctypes.POINTER(_vp)]),
        "pt_host_free": (_int, [_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status != PT_OK:
        msg = load().pt_last_error()
        raise PTError(status, msg.decode() if msg else "unknown error")


def ptr(a: np.ndarray, t=_vp):
    if a is None:
        return None
    if t is _vp:
        return ctypes.c_void_p(a.ctypes.data)
    return a.ctypes.data_as(t)


class DeviceBuffer:
    """A device array owned by a context: shape, dtype and a device pointer.
    Accepted wherever the bindings take device data (on_device = 1)."""

    def __init__(self, ctx, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(load().pt_buffer_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self._owner = True

    def view(self, shape):
        """A non-owning buffer over the first prod(shape) elements (C order)."""
        v = object.__new__(DeviceBuffer)
        v.ctx, v.dtype = self.ctx, self.dtype
        v.shape = tuple(int(s) for s in shape)
        v.nbytes = int(np.prod(v.shape)) * v.dtype.itemsize
        if v.nbytes > self.nbytes:
            raise ValueError("view larger than the buffer")
        v.ptr = self.ptr
        v._owner = False
        return v

    @classmethod
    def from_host(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(ctx, arr.shape, arr.dtype)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(load().pt_memcpy(self.ctx.h, ctypes.c_void_p(self.ptr), ptr(arr), self.nbytes, 0))

    def upload_columns(self, src, lo):
        """self (shape lead + (w,)) = src[..., lo:lo + w] for a C-contiguous
        host array src of shape lead + (n,): one strided copy by the copy
        engine, which reads page-locked memory (is_pinned) in place."""
        if not (isinstance(src, np.ndarray) and src.flags.c_contiguous and src.dtype == self.dtype):
            raise ValueError("upload_columns: src must be a C-contiguous array of the buffer's dtype")
        w, n, isz = self.shape[-1], src.shape[-1], self.dtype.itemsize
        if tuple(src.shape[:-1]) != tuple(self.shape[:-1]) or not 0 <= lo <= n - w:
            raise ValueError("upload_columns: columns out of range or shapes differ")
        rows = int(np.prod(self.shape[:-1]))
        check(load().pt_memcpy_2d(self.ctx.h, ctypes.c_void_p(self.ptr), w * isz,
                                  ctypes.c_void_p(src.ctypes.data + lo * isz), n * isz, w * isz, rows, 0))

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(load().pt_memcpy(self.ctx.h, ptr(out), ctypes.c_void_p(self.ptr), self.nbytes, 1))
        return out

    def row(self, i):
        """Pointer to row i of a C-ordered buffer (for per-component arrays)."""
        stride = int(np.prod(self.shape[1:])) * self.dtype.itemsize
        return ctypes.c_void_p(self.ptr + i * stride)

    def free(self):
        if self.ptr and self._owner:
            check(load().pt_buffer_free(self.ctx.h, ctypes.c_void_p(self.ptr), self.nbytes))
        self.ptr = None


_pinned = {}  # start address -> bytes of each live PinnedArray allocation


def is_pinned(arr):
    """True when all of arr's memory lies inside one live PinnedArray
    allocation (page-locked, readable by the device's copy engine)."""
    if not isinstance(arr, np.ndarray) or arr.size == 0:
        return False
    from numpy.lib.array_utils import byte_bounds
    lo, hi = byte_bounds(arr)
    return any(p <= lo and hi <= p + n for p, n in list(_pinned.items()))


def _host_free(ptr_value):
    _pinned.pop(ptr_value, None)
    try:
        load().pt_host_free(ctypes.c_void_p(ptr_value))
    except Exception:  # interpreter shutdown: the process releases it anyway
        pass


class PinnedArray:
    """A numpy view of page-locked host memory (pt_host_alloc).  The memory
    lives as long as any numpy view of it: it is released when the last view
    (``array``, slices of it, arrays handed out) is gone."""

    def __init__(self, shape, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(s) for s in shape)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(load().pt_host_alloc(nbytes, ctypes.byref(p)))
        self.ptr = p.value
        _pinned[self.ptr] = max(nbytes, 1)
        buf = (ctypes.c_char * max(nbytes, 1)).from_address(self.ptr)
        weakref.finalize(buf, _host_free, self.ptr)  # every view keeps buf alive
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    def free(self):
        """Drop this handle's view (the memory goes with the last view)."""
        self.array = None
        self.ptr = None


def exported_symbols():
    """Names of the C-ABI symbols present in the library (no GPU needed)."""
    L = ctypes.CDLL(LIB_PATH)
    return [n for n in EXPORTS if hasattr(L, n)]
