"""ctypes binding of libpt_h5.so (include/pt_h5.h): HDF5 files in the
reference's layouts.

HDF5.jl stores a Julia (a, b, c) array column-major, so the file's dataspace
is (c, b, a); numpy reads the same bytes C-ordered as shape (c, b, a).  This
module speaks numpy shapes throughout (a phase-space cache is (nsave, 3, n),
a mu cache (nsave, n)) and reverses them into the Julia-order dims the C ABI
takes.  Assignment follows HDF5.jl's `file[name] = value`:

    Python value                      stored as (HDF5.jl source form)
    float / int                       true scalar            (Int(x))
    np.array([x])                     length-1 vector        (Float64[x], Int[x])
    numpy array (f64, f32, i64, i32)  array, numpy shape     (Julia array, dims reversed)
    str                               scalar UTF-8 string    (string(x))
    list/tuple of str                 string vector          (Vector{String})
    bool                              Bool enum              (true / false)

The library must be built (python -m particle_transport_amd.build); there is
no fallback.
"""
from __future__ import annotations

import ctypes
import os
from numbers import Integral, Real

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpt_h5.so")

F64, F32, I64, I32 = 0, 1, 2, 3
NUMERIC, STRING, BOOL = 0, 1, 2
MAX_RANK = 8
_NP = {F64: np.float64, F32: np.float32, I64: np.int64, I32: np.int32}
_CODE = {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.int64): I64, np.dtype(np.int32): I32}

EXPORTS = ("pt_h5_last_error", "pt_h5_create", "pt_h5_open", "pt_h5_close", "pt_h5_flush", "pt_h5_exists",
           "pt_h5_create_group", "pt_h5_create_dataset", "pt_h5_write_box", "pt_h5_read_box", "pt_h5_write_array",
           "pt_h5_write_string", "pt_h5_write_strings", "pt_h5_write_bool", "pt_h5_info", "pt_h5_read_array",
           "pt_h5_read_string")

_lib = None
_vp, _cp, _i64 = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)


class H5Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise H5Error(f"{LIB_PATH} is missing: build it with `python -m particle_transport_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.pt_h5_last_error.restype = _cp
    L.pt_h5_last_error.argtypes = []
    sig = {
        "pt_h5_create": [_cp, ctypes.POINTER(_vp)],
        "pt_h5_open": [_cp, ctypes.c_int, ctypes.POINTER(_vp)],
        "pt_h5_close": [_vp],
        "pt_h5_flush": [_vp],
        "pt_h5_exists": [_vp, _cp, _ip],
        "pt_h5_create_group": [_vp, _cp],
        "pt_h5_create_dataset": [_vp, _cp, ctypes.c_int, ctypes.c_int, _i64p],
        "pt_h5_write_box": [_vp, _cp, ctypes.c_int, _i64p, _i64p, _vp],
        "pt_h5_read_box": [_vp, _cp, ctypes.c_int, _i64p, _i64p, _vp],
        "pt_h5_write_array": [_vp, _cp, ctypes.c_int, ctypes.c_int, _i64p, _vp],
        "pt_h5_write_string": [_vp, _cp, _cp],
        "pt_h5_write_strings": [_vp, _cp, _i64, ctypes.POINTER(_cp)],
        "pt_h5_write_bool": [_vp, _cp, ctypes.c_int],
        "pt_h5_info": [_vp, _cp, _ip, _ip, _ip, _i64p],
        "pt_h5_read_array": [_vp, _cp, ctypes.c_int, _vp, _i64],
        "pt_h5_read_string": [_vp, _cp, _i64, ctypes.c_char_p, _i64, _i64p],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def _check(status):
    if status != 0:
        raise H5Error(load().pt_h5_last_error().decode(errors="replace"))


def _julia(shape):
    """numpy shape -> Julia-order dims as a ctypes int64 array."""
    dims = [int(d) for d in reversed(tuple(shape))]
    if len(dims) > MAX_RANK:
        raise H5Error(f"rank {len(dims)} exceeds {MAX_RANK}")
    return (_i64 * max(1, len(dims)))(*dims) if dims else None


def _name(name):
    return str(name).encode()


class H5Dataset:
    """A dataset read lazily: shape/dtype in numpy order; ds[a:b] reads
    samples a..b-1 of the slowest axis (the Julia last axis); read_box and
    read_columns read hyperslabs."""

    def __init__(self, file, name):
        self.file, self.name = file, name
        kind, dtype, shape = file.info(name)
        if kind != NUMERIC:
            raise H5Error(f"'{name}' is not a numeric dataset")
        self.dtype = np.dtype(_NP[dtype])
        self.shape = shape
        self.ndim = len(shape)

    def __len__(self):
        return self.shape[0]

    def __array__(self, dtype=None, copy=None):
        a = self[...]
        return a if dtype is None else a.astype(dtype)

    def read_box(self, start, count, dtype=None, out=None):
        """numpy-order start/count; returns a C-ordered array of shape count
        (`out`, a C-contiguous array of that shape, when given)."""
        dt = np.dtype(dtype or (out.dtype if out is not None else self.dtype))
        shape = tuple(int(c) for c in count)
        if out is None:
            out = np.empty(shape, dtype=dt)
        elif out.shape != shape or out.dtype != dt or not out.flags.c_contiguous:
            raise H5Error("read_box: out must be a C-contiguous array of the box's shape and dtype")
        _check(load().pt_h5_read_box(self.file._h, _name(self.name), _CODE[dt], _julia(start), _julia(count),
                                     out.ctypes.data_as(_vp)))
        return out

    def __getitem__(self, key):
        if key is Ellipsis:
            return self.read_box((0,) * self.ndim, self.shape)
        if not isinstance(key, slice) or key.step not in (None, 1):
            raise H5Error("datasets read whole (ds[...]) or as a contiguous range of the first numpy axis")
        a, b, _ = key.indices(self.shape[0])
        b = max(a, b)
        return self.read_box((a,) + (0,) * (self.ndim - 1), (b - a,) + tuple(self.shape[1:]))

    def read_columns(self, cols, dtype=None):
        """Gather particles `cols` (0-based, along the last numpy axis = the
        Julia first axis) for every other index, one hyperslab per run of
        consecutive indices, as read_particle_batch does
        (compute_delta_mu2_dmumu_full.jl:411-437)."""
        return _read_columns(self, cols, dtype)


_COPY_THREADS = 8
_copy_pool = None


def _copy_rows(dst, src):
    """dst[...] = src (same shape), the rows of the first axis split over a
    few threads: a box of a cadence cache is many short strided rows, which
    one thread copies at a fraction of the memory bandwidth."""
    global _copy_pool
    n = dst.shape[0] if dst.ndim else 0
    k = min(_COPY_THREADS, n)
    if k < 2 or dst.nbytes < (8 << 20):
        np.copyto(dst, src, casting="same_kind")
        return
    if _copy_pool is None:
        from concurrent.futures import ThreadPoolExecutor
        _copy_pool = ThreadPoolExecutor(_COPY_THREADS)
    b = np.linspace(0, n, k + 1).astype(int)
    list(_copy_pool.map(lambda r: np.copyto(dst[r[0]:r[1]], src[r[0]:r[1]], casting="same_kind"),
                        zip(b[:-1], b[1:])))


def _reusable(count, dtype):
    """A flat host buffer of `count` elements, page-locked when the library
    can allocate it so (no device present: pageable)."""
    try:
        from ._lib import PinnedArray, PTError
    except ImportError:  # the HIP library is not built: pageable
        return np.empty(max(1, count), dtype=dtype)
    try:
        return PinnedArray((max(1, count),), dtype).array
    except (PTError, OSError):
        return np.empty(max(1, count), dtype=dtype)


def _read_columns(ds, cols, dtype=None):
    cols = np.asarray(cols, dtype=np.int64)
    if cols.size == 0:
        raise H5Error("Cannot read an empty particle batch.")
    if cols.min() < 0 or cols.max() >= ds.shape[-1]:
        raise H5Error("particle index out of range")
    dt = np.dtype(dtype or ds.dtype)
    out = np.empty(tuple(ds.shape[:-1]) + (len(cols),), dtype=dt)
    breaks = np.flatnonzero(np.diff(cols) != 1) + 1
    starts = np.concatenate([[0], breaks])
    stops = np.concatenate([breaks, [len(cols)]])
    lead = (0,) * (ds.ndim - 1)
    for s, e in zip(starts, stops):
        out[..., s:e] = ds.read_box(lead + (int(cols[s]),), tuple(ds.shape[:-1]) + (int(e - s),), dt)
    return out


class ArrayDataset:
    """A numpy array (or memory map) behind H5Dataset's reading interface,
    so in-memory and file-backed caches stream the same way."""

    def __init__(self, array):
        self.a = array
        self.shape = tuple(array.shape)
        self.dtype = np.dtype(array.dtype)
        self.ndim = len(self.shape)

    def read_box(self, start, count, dtype=None, out=None):
        sl = tuple(slice(int(s), int(s) + int(c)) for s, c in zip(start, count))
        src = self.a[sl]
        if out is None:
            out = np.empty(src.shape, dtype=dtype or self.dtype)
        elif out.shape != src.shape or not out.flags.c_contiguous:
            raise H5Error("read_box: out must be a C-contiguous array of the box's shape")
        _copy_rows(out, src)
        return out

    def read_columns(self, cols, dtype=None):
        return _read_columns(self, cols, dtype)


class H5File:
    """h5open(path, mode) for mode "r", "r+" or "w" (truncate)."""

    def __init__(self, path, mode="r"):
        L = load()
        self.path = str(path)
        h = _vp()
        if mode == "w":
            _check(L.pt_h5_create(self.path.encode(), ctypes.byref(h)))
        elif mode in ("r", "r+"):
            _check(L.pt_h5_open(self.path.encode(), int(mode == "r+"), ctypes.byref(h)))
        else:
            raise H5Error(f"unknown mode {mode!r}")
        self._h = h

    def close(self):
        if self._h:
            h, self._h = self._h, None
            _check(load().pt_h5_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flush(self):
        _check(load().pt_h5_flush(self._h))

    def __contains__(self, name):
        out = ctypes.c_int()
        _check(load().pt_h5_exists(self._h, _name(name), ctypes.byref(out)))
        return bool(out.value)

    def create_group(self, name):
        _check(load().pt_h5_create_group(self._h, _name(name)))

    def create_dataset(self, name, shape, dtype):
        """create_dataset(file, name, T, dims) with numpy shape = reversed dims."""
        dt = np.dtype(dtype)
        if len(shape) < 1:
            raise H5Error("create_dataset needs rank >= 1")
        _check(load().pt_h5_create_dataset(self._h, _name(name), _CODE[dt], len(shape), _julia(shape)))
        return H5Dataset(self, name)

    def write_box(self, name, start, array):
        a = np.ascontiguousarray(array)
        if a.dtype not in _CODE:
            a = a.astype(np.float64)
        _check(load().pt_h5_write_box(self._h, _name(name), _CODE[a.dtype], _julia(start), _julia(a.shape),
                                      a.ctypes.data_as(_vp)))

    def write_slab(self, name, first, array):
        """dataset[first:first + len(array)] = array along the first numpy
        axis (cadence samples), every other index whole."""
        a = np.ascontiguousarray(array)
        self.write_box(name, (int(first),) + (0,) * (a.ndim - 1), a)

    def __setitem__(self, name, value):
        L = load()
        if isinstance(value, (bool, np.bool_)):
            _check(L.pt_h5_write_bool(self._h, _name(name), int(bool(value))))
        elif isinstance(value, str):
            _check(L.pt_h5_write_string(self._h, _name(name), value.encode()))
        elif isinstance(value, (list, tuple)) and all(isinstance(v, str) for v in value) and len(value):
            arr = (_cp * len(value))(*[v.encode() for v in value])
            _check(L.pt_h5_write_strings(self._h, _name(name), len(value), arr))
        else:
            if isinstance(value, Integral):
                a = np.array(int(value), dtype=np.int64)
            elif isinstance(value, Real) and not isinstance(value, np.ndarray):
                a = np.array(float(value), dtype=np.float64)
            else:
                a = np.asarray(value)
                if a.dtype == np.bool_:
                    a = a.astype(np.int64)
                elif a.dtype.kind in "iu" and a.dtype not in _CODE:
                    a = a.astype(np.int64)
                elif a.dtype.kind == "f" and a.dtype not in _CODE:
                    a = a.astype(np.float64)
                elif a.dtype not in _CODE:
                    raise H5Error(f"cannot store {a.dtype} in '{name}'")
                a = np.ascontiguousarray(a)
            _check(L.pt_h5_write_array(self._h, _name(name), _CODE[a.dtype], a.ndim, _julia(a.shape),
                                       a.ctypes.data_as(_vp)))

    def info(self, name):
        kind, dtype, rank = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dims = (_i64 * MAX_RANK)()
        _check(load().pt_h5_info(self._h, _name(name), ctypes.byref(kind), ctypes.byref(dtype), ctypes.byref(rank),
                                 dims))
        return kind.value, dtype.value, tuple(int(dims[k]) for k in reversed(range(rank.value)))

    def dataset(self, name):
        return H5Dataset(self, name)

    def read_string(self, name, index=0):
        n = ctypes.c_int64()
        buf = ctypes.create_string_buffer(256)
        _check(load().pt_h5_read_string(self._h, _name(name), index, buf, len(buf), ctypes.byref(n)))
        if n.value >= len(buf):
            buf = ctypes.create_string_buffer(n.value + 1)
            _check(load().pt_h5_read_string(self._h, _name(name), index, buf, len(buf), ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    def __getitem__(self, name):
        """read(file[name]): numeric arrays in numpy shape, true scalars as
        Python numbers, strings as str (string vectors as lists), Bool as
        bool."""
        kind, dtype, shape = self.info(name)
        if kind == STRING:
            if len(shape) == 0:
                return self.read_string(name)
            return [self.read_string(name, k) for k in range(int(np.prod(shape)))]
        if kind == BOOL:
            out = np.empty(shape, dtype=np.int64)
            _check(load().pt_h5_read_array(self._h, _name(name), I64, out.ctypes.data_as(_vp), max(1, out.size)))
            return bool(out) if len(shape) == 0 else out.astype(bool)
        out = np.empty(shape, dtype=_NP[dtype])
        _check(load().pt_h5_read_array(self._h, _name(name), dtype, out.ctypes.data_as(_vp), max(1, out.size)))
        return out.item() if len(shape) == 0 else out


def scalar_value_or(file, name, default):
    """h5_scalar_value_or (multimode_multienergy_cache_pipeline.jl:1041-1049):
    a length-1 vector or a true scalar."""
    if name not in file:
        return default
    v = file[name]
    if isinstance(v, np.ndarray):
        return default if v.size == 0 else v.reshape(-1)[0].item()
    if isinstance(v, list):
        return default if not v else v[0]
    return v


def scalar_bool_or(file, name, default):
    """h5_scalar_bool_or (:1072-1080)."""
    v = scalar_value_or(file, name, default)
    if isinstance(v, (bool, np.bool_)):
        return bool(v)
    if isinstance(v, Integral):
        return v != 0
    s = str(v).strip().lower()
    if s in ("true", "yes", "1"):
        return True
    if s in ("false", "no", "0"):
        return False
    return default


def iter_particle_groups(datasets, pidx0, chunk, budget_bytes=256 << 20, prefetch=True, reuse=False):
    """Stream the selected particles of cadence datasets (numpy shapes
    (nsave, n) or (nsave, 3, n); particles on the last axis) in groups of
    whole analysis chunks.

    pidx0: 0-based selected particle indices, in selection order; chunk: the
    analysis chunk size.  Consecutive chunks are grouped while their index
    span fits budget_bytes per dataset; a group is read as one hyperslab
    over that span, or, when even one chunk's span is larger (scattered
    selections), as the runs of its selected indices (read_particle_batch,
    compute_delta_mu2_dmumu_full.jl:411-437).  The next group is read on a
    worker thread while the caller processes the current one.

    Yields (arrays, chunks): one array per dataset holding the group's
    columns, and per analysis chunk (offset into pidx0, its column indices
    into those arrays).

    reuse=True: a group's arrays are views of two sets of buffers allocated
    once (page-locked when the host allows, so a device upload from them is
    a straight DMA) and used in turn: a caller must be done with a group's
    arrays when it asks for the group after next, as a caller that uploads
    or consumes each group before asking for the next one is."""
    from concurrent.futures import ThreadPoolExecutor

    pidx0 = np.asarray(pidx0, dtype=np.int64)
    per_col = max(int(np.prod(ds.shape[:-1])) * ds.dtype.itemsize for ds in datasets)
    plans = plan_particle_groups(pidx0, chunk, max(1, int(budget_bytes) // max(1, per_col)))

    bufs = None
    if reuse:
        span = max((hi - lo + 1 for kind, _, _, lo, hi in plans if kind == "span"), default=0)
        if span:
            bufs = [[_reusable(int(np.prod(ds.shape[:-1])) * span, ds.dtype) for ds in datasets] for _ in range(2)]

    def read(plan, slot=0):
        kind, g0, g1, lo, hi = plan
        if kind == "span":
            arrays = []
            for k, ds in enumerate(datasets):
                lead = tuple(ds.shape[:-1])
                count = lead + (hi - lo + 1,)
                out = None
                if bufs is not None:
                    out = bufs[slot][k][:int(np.prod(count))].reshape(count)
                arrays.append(ds.read_box((0,) * len(lead) + (lo,), count, out=out))
        else:
            arrays = [ds.read_columns(pidx0[g0:g1]) for ds in datasets]
        return arrays, plan_chunks(plan, pidx0, chunk)

    if not prefetch:
        for k, plan in enumerate(plans):
            yield read(plan, k & 1)
        return
    with ThreadPoolExecutor(max_workers=1) as ex:
        fut = ex.submit(read, plans[0], 0) if plans else None
        for k in range(len(plans)):
            cur = fut.result()
            fut = ex.submit(read, plans[k + 1], (k + 1) & 1) if k + 1 < len(plans) else None
            yield cur


def plan_particle_groups(pidx0, chunk, cap):
    """The groups of iter_particle_groups: consecutive whole chunks of pidx0
    grouped while their index span stays within cap columns and the indices
    ascend, as ("span", g0, g1, lo, hi) (chunks pidx0[g0:g1], columns lo..hi
    inclusive), else one chunk as ("runs", g0, g1, 0, 0)."""
    pidx0 = np.asarray(pidx0, dtype=np.int64)
    plans = []
    c0 = 0
    while c0 < len(pidx0):
        g0 = c0
        lo = int(pidx0[c0])
        hi = int(pidx0[min(c0 + chunk, len(pidx0)) - 1])
        c1 = min(c0 + chunk, len(pidx0))
        if hi - lo + 1 <= cap and np.all(np.diff(pidx0[g0:c1]) > 0):
            # extend by whole chunks while the span fits and indices ascend
            while c1 < len(pidx0):
                n1 = min(c1 + chunk, len(pidx0))
                nhi = int(pidx0[n1 - 1])
                if nhi - lo + 1 > cap or not np.all(np.diff(pidx0[c1 - 1:n1]) > 0):
                    break
                c1, hi = n1, nhi
            plans.append(("span", g0, c1, lo, hi))
        else:
            plans.append(("runs", g0, c1, 0, 0))
        c0 = c1
    return plans


def plan_chunks(plan, pidx0, chunk):
    """Per analysis chunk of a group: (offset into pidx0, its column indices
    into the group's arrays: columns lo..hi of a span, the gathered columns
    of runs)."""
    kind, g0, g1, lo, _ = plan
    if kind == "span":
        return [(a, pidx0[a:min(a + chunk, g1)] - lo) for a in range(g0, g1, chunk)]
    return [(a, np.arange(a - g0, min(a + chunk, g1) - g0, dtype=np.int64)) for a in range(g0, g1, chunk)]
