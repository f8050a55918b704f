"""Trajectory engine host side: the reference's L3 entry points over libpt_hip.

``run_gpu_ensemble`` (phase_space_gpu_runner.jl:476-633),
``run_total_field_gpu_ensemble`` (:635-639), ``run_mu_cache_ensemble``
(multimode_multienergy_cache_pipeline.jl:1380-1535) and
``resolve_total_field_time_reference`` (phase_space_gpu_runner.jl:458-470) keep
their names, config keys and returned fields.  The per-step CUDA.jl kernel
launches, per-step ``sum(dalive)`` and save-step downloads are replaced by
``Ensemble.advance`` (pt_ens_advance), which runs many steps per launch and
writes the cadence samples straight into the cache arrays.

Classes ``Context``, ``Field`` and ``Ensemble`` wrap the C-ABI handles.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from types import SimpleNamespace

import numpy as np

from . import _lib, cache_h5, h5
from ._lib import PT_F32, PT_F64, PUSHER, STORAGE, check, ptr
from .fields import build_uniform_coords, estimate_min_dx
from .injection import (M_P, Q_E, energy_to_gamma, energy_to_speed, init_particles, inject_params,
                        injection_mode, injection_position_mode, injection_position_tuple,
                        injection_position_unit, injection_rng)
from .time_units import (TIME_REFERENCE_MODE, PipelineError, campaign_time_reference, reference_Omega0,
                         require_explicit_campaign_time_reference, resolve_save_stride,
                         resolve_trajectory_time_grid, sampled_step_indices)


def _julia_layout_pointer(a, keep):
    """Pointer to a host (nx,ny,nz) array whose memory has x fastest.
    Returns (pointer, dtype_code, on_device, shape)."""
    arr = np.asarray(a)
    if arr.dtype not in (np.float32, np.float64):
        arr = arr.astype(np.float64)
    arr = np.asfortranarray(arr)
    shape = arr.shape
    arr = arr.transpose(2, 1, 0)
    keep.append(arr)
    assert arr.flags.c_contiguous
    return ptr(arr), (PT_F64 if arr.dtype == np.float64 else PT_F32), 0, shape


def _vec_pointer(v, keep, dtype=None):
    a = np.ascontiguousarray(v, dtype=dtype or (np.float64 if np.asarray(v).dtype == np.float64 else np.float32))
    keep.append(a)
    return ptr(a), (PT_F64 if a.dtype == np.float64 else PT_F32), 0


class Context:
    """A device and the library context (stream, scratch) on it."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.pt_ctx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            check(_lib.load().pt_ctx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def bytes_in_use(self) -> int:
        v = ctypes.c_int64()
        check(_lib.load().pt_ctx_bytes_in_use(self.h, ctypes.byref(v)))
        return v.value

    def device_memory(self):
        """(free, total) bytes of the device (pt_device_memory)."""
        f, t = ctypes.c_int64(), ctypes.c_int64()
        check(_lib.load().pt_device_memory(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def mean_abs_b(self, Bx, By, Bz) -> float:
        """reference_B0_T (time_units.jl:18-20) as an FP64 device reduction."""
        keep = []
        pb = [_julia_layout_pointer(b, keep) if np.ndim(b) == 3 else (*_vec_pointer(np.ravel(b), keep), None)
              for b in (Bx, By, Bz)]
        dt = {p[1] for p in pb}
        dev = {p[2] for p in pb}
        if len(dt) != 1 or len(dev) != 1:
            raise PipelineError("Bx, By, Bz must share dtype and device")
        n = int(np.prod(np.shape(Bx)))
        out = ctypes.c_double()
        check(_lib.load().pt_mean_abs_b(self.h, n, dt.pop(), dev.pop(), pb[0][0], pb[1][0], pb[2][0],
                                        ctypes.byref(out)))
        return out.value


class Field:
    """The six field components on the device in one storage layout.

    cell_table=True also builds the cell table (pt_field_cell_table: 192
    bytes per cell of precomputed FP32 polynomial coefficients) that the
    production pushers then read on every cell change; results are
    bit-identical with and without it."""

    def __init__(self, ctx: Context, comps, box_length_m: float, storage: str = "f64_planar", *,
                 cell_table: bool = False):
        keep = []
        ptrs, dts, devs, shapes = zip(*[_julia_layout_pointer(c, keep) for c in comps])
        if len(set(dts)) != 1 or len(set(devs)) != 1 or len(set(shapes)) != 1:
            raise PipelineError("field components must share shape, dtype and device")
        nx, ny, nz = shapes[0]
        arr = (ctypes.c_void_p * 6)(*ptrs)
        h = ctypes.c_void_p()
        check(_lib.load().pt_field_create(ctx.h, nx, ny, nz, float(box_length_m), STORAGE[storage], dts[0],
                                          devs[0], arr, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.shape = (nx, ny, nz)
        self.storage = storage
        self.box_length_m = float(box_length_m)
        self.cell_table = False
        if cell_table:
            self.set_cell_table(True)

    def set_cell_table(self, enable: bool):
        """Build (True) or free (False) the field's cell table."""
        check(_lib.load().pt_field_cell_table(self.h, 1 if enable else 0))
        self.cell_table = bool(enable)

    @property
    def B0(self) -> float:
        v = ctypes.c_double()
        check(_lib.load().pt_field_mean_abs_b(self.h, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            check(_lib.load().pt_field_destroy(self.h))
            self.h = None


class Ensemble:
    """Particle state on the device, advanced by one of the pushers."""

    def __init__(self, ctx: Context, field: Field, x, p, *, pusher: str = "ref_f64", periodic: bool = True):
        keep = []
        xs = [_vec_pointer(v, keep, np.float64) for v in x]
        ps = [_vec_pointer(v, keep, np.float64) for v in p]
        n = int(len(x[0]))
        dts = {v[1] for v in xs + ps}
        devs = {v[2] for v in xs + ps}
        if len(dts) != 1 or len(devs) != 1:
            raise PipelineError("state arrays must share dtype and device")
        xa = (ctypes.c_void_p * 3)(*[v[0] for v in xs])
        pa = (ctypes.c_void_p * 3)(*[v[0] for v in ps])
        h = ctypes.c_void_p()
        check(_lib.load().pt_ens_create(ctx.h, field.h, PUSHER[pusher], n, dts.pop(), devs.pop(), xa, pa,
                                        int(bool(periodic)), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.field = field
        self.n = n
        self.pusher = pusher

    @classmethod
    def injected(cls, ctx: Context, field: Field, n: int, params, *, pusher: str = "ref_f64",
                 periodic: bool = True):
        """pt_ens_create_injected: the initial state is drawn on the device
        (init_particles with per-particle Philox4x64-10 counters); params is
        injection.inject_params(...)."""
        h = ctypes.c_void_p()
        check(_lib.load().pt_ens_create_injected(ctx.h, field.h, PUSHER[pusher], int(n), ctypes.byref(params),
                                                 int(bool(periodic)), ctypes.byref(h)))
        self = cls.__new__(cls)
        self.h = h
        self.ctx = ctx
        self.field = field
        self.n = int(n)
        self.pusher = pusher
        return self

    def advance(self, field=None, *, dt, first_step, n_steps, total_steps, stride=1, out_kind="none",
                out_dtype=np.float32, charge=Q_E, mass=M_P, steps_per_launch=0, out_a=None, out_b=None,
                out_capacity=None, sort_every_steps=0):
        """pt_ens_advance.  With out_a/out_b None, host arrays are allocated:
        phase space -> (nsave, 3, n) each (Julia (n,3,nsave)); mu -> (nsave, n).
        _lib.DeviceBuffer outputs keep the cadence samples in HBM."""
        field = field or self.field
        kind = _lib.OUT[out_kind]
        nsv = _saves_in_range(first_step, n_steps, stride) if kind else 0
        code = PT_F64 if np.dtype(out_dtype) == np.float64 else PT_F32
        on_dev = 0
        if kind and nsv:
            if out_a is None:
                shape = (nsv, 3, self.n) if kind == 1 else (nsv, self.n)
                out_a = np.empty(shape, dtype=out_dtype)
                out_b = np.empty(shape, dtype=out_dtype) if kind == 1 else None
                cap = nsv
            else:
                cap = out_capacity if out_capacity is not None else nsv
            if isinstance(out_a, _lib.DeviceBuffer):
                on_dev = 1
                pa = ctypes.c_void_p(out_a.ptr)
                pb = ctypes.c_void_p(out_b.ptr) if out_b is not None else None
            else:
                pa, pb = ptr(out_a), (ptr(out_b) if out_b is not None else None)
        else:
            pa = pb = None
            cap = 0
        prm = _lib.AdvanceParams(float(dt), float(charge), float(mass), int(first_step), int(n_steps),
                                 int(total_steps), int(stride), int(kind), int(steps_per_launch),
                                 int(sort_every_steps))
        alive = np.zeros(max(int(n_steps), 1), dtype=np.int64)
        nsaved = ctypes.c_int64()
        check(_lib.load().pt_ens_advance(self.h, field.h, ctypes.byref(prm), pa, pb, code, on_dev, int(cap),
                                         ptr(alive, _lib._i64p), ctypes.byref(nsaved)))
        ms = ctypes.c_double()
        nl = ctypes.c_int64()
        check(_lib.load().pt_ens_last_timing(self.h, ctypes.byref(ms), ctypes.byref(nl)))
        return SimpleNamespace(out_a=out_a if nsv else None, out_b=out_b if nsv else None,
                               alive_counts=alive[:int(n_steps)], n_saved=nsaved.value, kernel_ms=ms.value,
                               launches=nl.value)

    def sort(self):
        """Morton re-sort of the particles by brick (pt_ens_sort)."""
        check(_lib.load().pt_ens_sort(self.h))

    def state(self):
        """(x (3,n) metres, p (3,n) kg m/s, alive (n,) bool) as float64 host arrays."""
        x = np.empty((3, self.n))
        p = np.empty((3, self.n))
        alive = np.empty(self.n, dtype=np.uint8)
        xp = (ctypes.POINTER(ctypes.c_double) * 3)(*[x[c].ctypes.data_as(_lib._dp) for c in range(3)])
        pp = (ctypes.POINTER(ctypes.c_double) * 3)(*[p[c].ctypes.data_as(_lib._dp) for c in range(3)])
        check(_lib.load().pt_ens_get_state(self.h, xp, pp, alive.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return x, p, alive.astype(bool)

    def close(self):
        if self.h:
            check(_lib.load().pt_ens_destroy(self.h))
            self.h = None


def _saves_in_range(first, count, stride):
    if count <= 0:
        return 0
    a, b = first - 1, first + count - 2
    lo = -(-a // stride)
    hi = b // stride
    return max(0, hi - lo + 1)


# --------------------------------------------------------------- caches ----

class CacheWriter:
    """Cadence cache in the reference's dataset layout (HDF5.jl writes a Julia
    (n,3,nsave) array, seen from C as (nsave,3,n); numpy uses the C view).

    output path ending in ".h5": an HDF5 file in the reference's cache layout
      (create_phase_space_writer / create_mu_cache_writer,
      phase_space_gpu_runner.jl:406-446, multimode_multienergy_cache_pipeline.jl:
      1312-1347), written through libpt_h5.so, one hyperslab per drained chunk;
    any other path: .npy memory maps next to a metadata.json;
    no path: in memory, page-locked so that device chunks are copied straight
      into them."""

    def __init__(self, path, datasets: dict, dtype=np.float32):
        self.path = path
        self.dtype = dtype
        self.data = {}
        self.meta = {}
        self.file = None
        self.pinned = not path
        if path and str(path).endswith(".h5"):
            os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
            self.file = h5.H5File(path, "w")
            for name, shape in datasets.items():
                self.data[name] = self.file.create_dataset(name, shape, dtype)
            return
        if path:
            os.makedirs(path, exist_ok=True)
        for name, shape in datasets.items():
            if path:
                self.data[name] = np.lib.format.open_memmap(os.path.join(path, name + ".npy"), mode="w+",
                                                            dtype=dtype, shape=shape)
            else:
                try:
                    self.data[name] = _lib.PinnedArray(shape, dtype).array
                except _lib.PTError:  # more than the host will page-lock: pageable + staging
                    self.pinned = False
                    self.data[name] = np.empty(shape, dtype=dtype)

    @property
    def is_h5(self):
        return self.file is not None or (bool(self.path) and str(self.path).endswith(".h5"))

    def __getitem__(self, k):
        return self.data[k]

    def write(self, name, offset, block):
        """Samples offset .. offset + len(block) - 1 of dataset `name`."""
        if self.file is not None:
            self.file.write_slab(name, offset, block)
        else:
            self.data[name][offset:offset + len(block)] = block

    def finalize(self, meta: dict, h5_meta=None):
        """Record the metadata; an HDF5 cache gets it written by h5_meta(file)
        in the reference's forms, is closed, and is reopened read-only so the
        datasets stay readable through this writer."""
        self.meta.update(meta)
        if self.file is not None:
            if h5_meta is not None:
                h5_meta(self.file)
            self.file.close()
            self.file = None
            reader = h5.H5File(self.path, "r")
            self.data = {k: reader.dataset(k) for k in self.data}
            self._reader = reader
        elif self.path:
            for a in self.data.values():
                a.flush()
            with open(os.path.join(self.path, "metadata.json"), "w") as f:
                json.dump({k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in self.meta.items()}, f,
                          default=str)

    def close(self):
        if self.file is not None:
            self.file.close()
            self.file = None
        r = getattr(self, "_reader", None)
        if r is not None:
            r.close()
            self._reader = None


def _precision_dtype(cfg):
    prec = str(cfg.get("precision", "Float64"))
    return np.float64 if prec in ("Float64", "float64") else np.float32


def default_engine(cfg):
    """Pusher and field storage for a config: precision Float64 keeps the
    reference's FP64 structure; Float32/BFloat16/Float16 use the production
    pusher with FP32/BF16/FP16 brick fields.  cfg["gpu_pusher"] and
    cfg["field_storage"] override."""
    prec = str(cfg.get("precision", "Float64"))
    pusher, storage = {"Float64": ("ref_f64", "f64_planar"), "Float32": ("fast", "f32_brick"),
                       "BFloat16": ("fast", "bf16_brick"), "Float16": ("fast", "f16_brick")}.get(
        prec, ("ref_f64", "f64_planar"))
    return cfg.get("gpu_pusher", pusher), cfg.get("field_storage", storage)


def resolve_total_field_time_reference(cfg, fields, *, energy_GeV, ctx=None):
    """phase_space_gpu_runner.jl:458-470."""
    mode = str(cfg.get("trajectory_mode", "unknown")).lower()
    if mode != "total":
        raise PipelineError("Trajectory mode '" + mode + "' requires an explicit campaign time reference derived "
                            "from the total-field dataset. The low-level runner must not derive B0_reference from "
                            "a decomposed or unknown field.")
    ref_mode = str(cfg.get("time_reference_mode", TIME_REFERENCE_MODE)).lower().replace("_", "-")
    if ref_mode != TIME_REFERENCE_MODE:
        raise PipelineError("Standalone total-field reference supports only " + TIME_REFERENCE_MODE + ".")
    own = ctx is None
    c = Context() if own else ctx
    try:
        B0 = c.mean_abs_b(fields[0], fields[1], fields[2])
    finally:
        if own:
            c.close()
    gamma0 = energy_to_gamma(energy_GeV, _precision_dtype(cfg))
    Omega0 = reference_Omega0(B0, gamma0, Q_E, M_P)
    src = str(cfg.get("trajectory_field_source_path", cfg.get("file", "")))
    return campaign_time_reference(B0, Omega0, source_mode="total", source_path=src,
                                   source_identity=str(cfg.get("trajectory_field_source_identity", src)),
                                   field_subset=cfg.get("field_subset"))


def _trajectory_setup(cfg, fields, energy_GeV, time_reference, caller):
    T = _precision_dtype(cfg)
    nx, ny, nz = np.shape(fields[0])
    axes = build_uniform_coords(cfg["box_length_pc"], nx, ny, nz)
    boundary = str(cfg.get("boundary", "periodic")).strip(":").lower()
    if boundary not in ("open", "periodic"):
        raise PipelineError("Unknown boundary mode: " + boundary + ". Use :open or :periodic.")
    gamma0 = energy_to_gamma(energy_GeV, T)
    time_reference = require_explicit_campaign_time_reference(str(cfg.get("trajectory_mode", "unknown")).lower(),
                                                              time_reference, caller=caller)
    Omega0 = T(time_reference.Omega0_reference_s_inv)
    v0 = energy_to_speed(energy_GeV, T)
    tt = resolve_trajectory_time_grid(cfg, Omega0, v0, estimate_min_dx(axes))
    st = resolve_save_stride(cfg, tt)
    dt = T(tt.dt_s)
    nsteps = tt.n_integration_steps
    t_s = dt * np.arange(nsteps, dtype=T)
    t_norm = t_s * Omega0
    t_gp = t_norm / T(2 * math.pi)
    save_idx = sampled_step_indices(nsteps, st.trajectory_time_stride)
    return SimpleNamespace(T=T, axes=axes, boundary=boundary, gamma0=gamma0, Omega0=Omega0, v0=v0, tt=tt, st=st,
                           dt=dt, nsteps=nsteps, t_s=t_s, t_norm=t_norm, t_gp=t_gp, save_idx=save_idx,
                           time_reference=time_reference, shape=(nx, ny, nz))


def _run_loop(cfg, fields, setup, out_kind, writer, ctx, field, initial_state, chunk_saves):
    """Shared step loop of run_gpu_ensemble / run_mu_cache_ensemble.

    Each pt_ens_advance call covers chunk_saves cadence samples and writes them
    into one of two device buffers; a second context (its own stream) copies
    the finished buffer into pinned host memory and hands it to the cache
    writer on a worker thread while the next call pushes into the other
    buffer, so the download and the file writes overlap the pusher (the
    reference downloads and writes synchronously at every save step,
    phase_space_gpu_runner.jl:578-590)."""
    from concurrent.futures import ThreadPoolExecutor

    s = setup
    n = int(cfg["n_particles"])
    on_device = initial_state is None and injection_rng(cfg) == "device"
    if initial_state is not None:
        x1, x2, x3, p1, p2, p3 = initial_state
    elif not on_device:
        x1, x2, x3, p1, p2, p3 = init_particles(cfg, s.axes, s.gamma0, s.v0, s.T, fields[0], fields[1], fields[2])
    pusher, storage = default_engine(cfg)
    own_ctx = ctx is None
    ctx = ctx or Context()
    own_field = field is None
    L = float(cfg["box_length_pc"]) * 3.085677581e16
    field = field or Field(ctx, fields, L, storage)
    periodic = s.boundary == "periodic"
    if on_device:
        ens = Ensemble.injected(ctx, field, n, inject_params(cfg, s.axes, s.gamma0, s.v0), pusher=pusher,
                                periodic=periodic)
    else:
        ens = Ensemble(ctx, field, (x1, x2, x3), (p1, p2, p3), pusher=pusher, periodic=periodic)
    alive_fraction = np.empty(s.nsteps)
    stride = s.st.trajectory_time_stride
    nsave = len(s.save_idx)
    out_T = writer.dtype
    shape = (chunk_saves, 3, n) if out_kind == "phase" else (chunk_saves, n)
    names = ("positions", "momenta") if out_kind == "phase" else ("mu",)
    dev = [[_lib.DeviceBuffer(ctx, shape, out_T) for _ in names] for _ in range(2)]
    # a page-locked in-memory cache takes the copies directly; a file-backed
    # one goes through page-locked staging buffers
    pinned = None if writer.pinned else [[_lib.PinnedArray(shape, out_T) for _ in names] for _ in range(2)]
    copy_ctx = Context(ctx.device)
    pool = ThreadPoolExecutor(max_workers=1)
    busy = [None, None]
    kernel_ms = 0.0

    io = {"download_s": 0.0, "write_s": 0.0}

    def drain(buf, offset, count):
        for k, name in enumerate(names):
            nbytes = count * int(np.prod(shape[1:])) * np.dtype(out_T).itemsize
            t0 = time.perf_counter()
            if pinned is None:
                dst = writer[name][offset:offset + count]
                check(_lib.load().pt_memcpy(copy_ctx.h, ctypes.c_void_p(dst.ctypes.data),
                                            ctypes.c_void_p(dev[buf][k].ptr), nbytes, 1))
                io["download_s"] += time.perf_counter() - t0
            else:
                check(_lib.load().pt_memcpy(copy_ctx.h, ctypes.c_void_p(pinned[buf][k].ptr),
                                            ctypes.c_void_p(dev[buf][k].ptr), nbytes, 1))
                t1 = time.perf_counter()
                writer.write(name, offset, pinned[buf][k].array[:count])
                io["download_s"] += t1 - t0
                io["write_s"] += time.perf_counter() - t1

    try:
        t0 = time.perf_counter()
        si, save_ptr, it = 1, 0, 0
        progress = int(cfg.get("progress_every", 5000))
        while si <= s.nsteps:
            count = min(stride * chunk_saves, s.nsteps - si + 1)
            buf = it % 2
            if busy[buf] is not None:
                busy[buf].result()  # the copy out of this buffer must finish before it is rewritten
            r = ens.advance(field, dt=float(s.dt), first_step=si, n_steps=count, total_steps=s.nsteps,
                            stride=stride, out_kind=out_kind, out_dtype=out_T, out_a=dev[buf][0],
                            out_b=dev[buf][1] if out_kind == "phase" else None, out_capacity=chunk_saves,
                            steps_per_launch=int(cfg.get("steps_per_launch", 0)),
                            sort_every_steps=int(cfg.get("sort_every_steps", 0)))
            alive_fraction[si - 1:si - 1 + count] = r.alive_counts / float(n)
            if r.n_saved:
                busy[buf] = pool.submit(drain, buf, save_ptr, r.n_saved)
            save_ptr += r.n_saved
            kernel_ms += r.kernel_ms
            if progress and (si - 1) // progress != (si - 1 + count) // progress:
                print(f"Energy step {si + count - 1}/{s.nsteps}")
            si += count
            it += 1
        for f in busy:
            if f is not None:
                f.result()
        elapsed = time.perf_counter() - t0
        assert save_ptr == nsave
    finally:
        pool.shutdown(wait=True)
        for group in dev:
            for d in group:
                d.free()
        for group in pinned or ():
            for h in group:
                h.free()
        copy_ctx.close()
        ens.close()
        if own_field:
            field.close()
        if own_ctx:
            ctx.close()
    return alive_fraction, elapsed, kernel_ms, io


def _common_meta(cfg, s, energy_GeV, alive_fraction):
    idx = s.save_idx - 1
    tr = s.time_reference
    return dict(
        t_norm=s.t_norm[idx].astype(np.float64), t_s=s.t_s[idx].astype(np.float64),
        t_gyroperiods=s.t_gp[idx].astype(np.float64), alive_fraction=alive_fraction[idx],
        energy_GeV=float(energy_GeV), n_particles=int(cfg["n_particles"]),
        Omega0=tr.Omega0_reference_s_inv, B0_T=tr.B0_reference_T, dt_s=s.tt.dt_s,
        dt_gyroperiods=s.tt.dt_gyroperiods, integration_step_count=s.tt.n_integration_steps,
        analysis_sample_count=len(idx), trajectory_time_stride=s.st.trajectory_time_stride,
        exact_final_state_stored=False, boundary_mode=s.boundary,
        trajectory_mode=str(cfg.get("trajectory_mode", "unknown")), **injection_metadata(cfg))


def injection_metadata(cfg):
    """write_injection_metadata! (phase_space_gpu_runner.jl:136-143), plus where
    the draws were made (injection_rng)."""
    return dict(injection_mode=injection_mode(cfg), injection_mu0=[float(cfg.get("injection_mu0", 0.0))],
                injection_position_mode=injection_position_mode(cfg),
                injection_position=[float(v) for v in injection_position_tuple(cfg, np.float64)],
                injection_position_unit=injection_position_unit(cfg), injection_rng=injection_rng(cfg))


def run_gpu_ensemble(cfg, fields, *, energy_GeV, time_reference, ctx=None, field=None, initial_state=None,
                     output_path=None, chunk_saves=64):
    """phase_space_gpu_runner.jl:476-633: phase-space cache at cadence.
    Returns the reference's NamedTuple fields plus ``cache`` (the writer)."""
    s = _trajectory_setup(cfg, fields, energy_GeV, time_reference, "low-level runner")
    out_T = np.float64 if str(cfg.get("trajectory_output_precision", "Float32")) == "Float64" else np.float32
    n = int(cfg["n_particles"])
    writer = CacheWriter(output_path, {"positions": (len(s.save_idx), 3, n), "momenta": (len(s.save_idx), 3, n)},
                         out_T)
    try:
        alive, elapsed, kms, io = _run_loop(cfg, fields, s, "phase", writer, ctx, field, initial_state,
                                            chunk_saves)
    except BaseException:
        writer.close()
        raise
    meta = _common_meta(cfg, s, energy_GeV, alive)
    meta.update(trajectory_output_precision="Float64" if out_T == np.float64 else "Float32", position_unit="m",
                momentum_unit="kg*m/s")
    writer.finalize(meta, lambda f: cache_h5.finalize_cache(f, cfg, s, energy_GeV, alive, "phase", out_T))
    return SimpleNamespace(
        t_norm=s.t_norm.astype(np.float64), t_s=s.t_s.astype(np.float64), t_gyroperiods=s.t_gp.astype(np.float64),
        alive_fraction=alive, nsave=len(s.save_idx), phase_space_path=output_path,
        phase_space_gib=6 * n * len(s.save_idx) * np.dtype(out_T).itemsize / 1024.0 ** 3, elapsed=elapsed,
        kernel_ms=kms, io_s=io, Omega0=float(s.Omega0), dt=float(s.dt), trajectory_time=s.tt, save_time=s.st,
        B0=float(s.time_reference.B0_reference_T), time_reference=s.time_reference, cache=writer)


def run_total_field_gpu_ensemble(cfg, fields, *, energy_GeV, **kw):
    """phase_space_gpu_runner.jl:635-639."""
    total = dict(cfg, trajectory_mode="total")
    tr = resolve_total_field_time_reference(total, fields, energy_GeV=energy_GeV, ctx=kw.get("ctx"))
    return run_gpu_ensemble(total, fields, energy_GeV=energy_GeV, time_reference=tr, **kw)


def run_mu_cache_ensemble(cfg, fields, output_path=None, *, energy_GeV, time_reference, ctx=None, field=None,
                          initial_state=None, chunk_saves=256):
    """multimode_multienergy_cache_pipeline.jl:1380-1535: mu cache at cadence,
    with mu fused into the push (reconstruct_mu_step_kernel!, :1349-1374)."""
    s = _trajectory_setup(cfg, fields, energy_GeV, time_reference, "compact mu cache runner")
    out_T = np.float64 if str(cfg.get("mu_cache_output_precision", "Float32")) == "Float64" else np.float32
    n = int(cfg["n_particles"])
    writer = CacheWriter(output_path, {"mu": (len(s.save_idx), n)}, out_T)
    try:
        alive, elapsed, kms, io = _run_loop(cfg, fields, s, "mu", writer, ctx, field, initial_state,
                                            chunk_saves)
    except BaseException:
        writer.close()
        raise
    meta = _common_meta(cfg, s, energy_GeV, alive)
    meta.update(cache_mode="mu", cache_output_precision="Float64" if out_T == np.float64 else "Float32",
                mu_unit="dimensionless")
    writer.finalize(meta, lambda f: cache_h5.finalize_cache(f, cfg, s, energy_GeV, alive, "mu", out_T))
    return SimpleNamespace(cache_path=output_path, cache_gib=n * len(s.save_idx) * np.dtype(out_T).itemsize / 1024.0 ** 3,
                           elapsed=elapsed, kernel_ms=kms, io_s=io, Omega0=float(s.Omega0), dt=float(s.dt),
                           trajectory_time=s.tt, save_time=s.st, B0=float(s.time_reference.B0_reference_T),
                           time_reference=s.time_reference, cache=writer, alive_fraction=alive,
                           t_s=s.t_s, t_norm=s.t_norm, t_gyroperiods=s.t_gp)


def run_total_field_mu_cache_ensemble(cfg, fields, output_path=None, *, energy_GeV, **kw):
    """multimode_multienergy_cache_pipeline.jl:1537-1541."""
    total = dict(cfg, trajectory_mode="total")
    tr = resolve_total_field_time_reference(total, fields, energy_GeV=energy_GeV, ctx=kw.get("ctx"))
    return run_mu_cache_ensemble(total, fields, output_path, energy_GeV=energy_GeV, time_reference=tr, **kw)
