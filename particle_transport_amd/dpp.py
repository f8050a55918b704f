"""D_pp(tau) and kinetic-energy snapshots, host side over the GPU kernels.

Restates compute_dpp_full.jl:232-330 and the product assembly of run_dpp_full
(:456-541).  ``process_momenta_chunk_dpp`` (process_momenta_chunk_dpp_cpu!,
:247-288) and ``fill_energy_snapshots`` (:290-299) run on the GPU through
pt_dpp_accumulate / pt_energy_snapshots.  Momenta use the phase-space cache
layout Julia (n,3,nsave), i.e. numpy (nsave, 3, n), float32.
"""
from __future__ import annotations

import ctypes
import math
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import check, ptr
from .dmumu import build_particle_indices, device_particle_groups, gpu_batch_particles
from .injection import C_LIGHT, GEV_TO_J, M_P
from .julia_compat import range_int_length, round_int, unique_stable
from .time_units import PipelineError, resolve_lag_grid, t_gyroperiods_from_axes, validate_time_axes


def momentum_magnitude(px, py, pz) -> float:
    """compute_dpp_full.jl:232-234 (FP64)."""
    px, py, pz = float(px), float(py), float(pz)
    return math.sqrt(px * px + py * py + pz * pz)


def kinetic_energy_gev_from_pmag(pmag) -> float:
    """compute_dpp_full.jl:236-240."""
    if not math.isfinite(pmag):
        return math.nan
    gamma = math.sqrt(1.0 + (float(pmag) / (M_P * C_LIGHT)) ** 2)
    return (gamma - 1.0) * M_P * C_LIGHT ** 2 / GEV_TO_J


def selected_energy_snapshot_indices(nsteps: int, count: int):
    """compute_dpp_full.jl:242-245, 1-based like the reference."""
    k = min(max(1, int(count)), int(nsteps))
    return unique_stable([round_int(v) for v in range_int_length(1, int(nsteps), k)])


def new_accumulators(n_lags: int):
    return SimpleNamespace(counts=np.zeros(n_lags, np.int64), sum_delta_p=np.zeros(n_lags),
                           sum_delta_p2=np.zeros(n_lags), sum_delta_p_norm=np.zeros(n_lags),
                           sum_delta_p_norm2=np.zeros(n_lags))


def _mom_pointer(momenta):
    if isinstance(momenta, _lib.DeviceBuffer):
        return momenta, ctypes.cast(ctypes.c_void_p(momenta.ptr), _lib._fp), 1, momenta.shape
    a = np.ascontiguousarray(momenta, dtype=np.float32)
    return a, a.ctypes.data_as(_lib._fp), 0, a.shape


def process_momenta_chunk_dpp(ctx, momenta, lag_steps, p0, acc, particle_index=None):
    """process_momenta_chunk_dpp_cpu! on the GPU.  particle_index: 0-based
    columns of momenta to use (default: all)."""
    keep, p, dev, shape = _mom_pointer(momenta)
    nsave, three, total = shape
    if three != 3:
        raise PipelineError("momenta dataset must have shape (particles, 3, steps).")
    idx = None if particle_index is None else np.ascontiguousarray(particle_index, dtype=np.int64)
    nsel = total if idx is None else len(idx)
    lags = np.ascontiguousarray(lag_steps, dtype=np.int64)
    check(_lib.load().pt_dpp_accumulate(ctx.h, p, dev, nsave, total, nsel,
                                        ptr(idx, _lib._i64p) if idx is not None else None, ptr(lags, _lib._i64p),
                                        len(lags), float(p0), ptr(acc.counts, _lib._i64p),
                                        ptr(acc.sum_delta_p, _lib._dp), ptr(acc.sum_delta_p2, _lib._dp),
                                        ptr(acc.sum_delta_p_norm, _lib._dp), ptr(acc.sum_delta_p_norm2, _lib._dp)))
    return acc


def fill_energy_snapshots(ctx, energy_snapshots, momenta, snapshot_indices, particle_offset, particle_index=None):
    """fill_energy_snapshots! on the GPU.  energy_snapshots: Fortran-ordered
    (k, N_sel) float64; snapshot_indices 1-based; particle_offset 1-based."""
    keep, p, dev, shape = _mom_pointer(momenta)
    nsave, _, total = shape
    idx = None if particle_index is None else np.ascontiguousarray(particle_index, dtype=np.int64)
    nsel = total if idx is None else len(idx)
    snap = np.ascontiguousarray(np.asarray(snapshot_indices, dtype=np.int64) - 1)
    if not energy_snapshots.flags.f_contiguous or energy_snapshots.dtype != np.float64:
        raise PipelineError("energy_snapshots must be a Fortran-ordered float64 array")
    k = energy_snapshots.shape[0]
    if k != len(snap) or particle_offset - 1 + nsel > energy_snapshots.shape[1]:
        raise PipelineError("energy_snapshots shape does not match the snapshot/particle counts")
    check(_lib.load().pt_energy_snapshots(ctx.h, p, dev, nsave, total, nsel,
                                          ptr(idx, _lib._i64p) if idx is not None else None, ptr(snap, _lib._i64p),
                                          k, energy_snapshots.ctypes.data_as(_lib._dp), int(particle_offset) - 1))
    return energy_snapshots


def compute_dpp_arrays(acc, tau_s, tau_norm):
    """compute_dpp_full.jl:301-330."""
    n = len(acc.counts)
    out = {k: np.full(n, np.nan) for k in ("mean_delta_p", "mean_delta_p2", "mean_delta_p_norm",
                                            "mean_delta_p_norm2", "dpp_raw_per_s", "dpp_centered_per_s",
                                            "dpp_raw_norm", "dpp_centered_norm")}
    for li in range(n):
        c = acc.counts[li]
        if not c > 0:
            continue
        dm = acc.sum_delta_p[li] / c
        d2 = acc.sum_delta_p2[li] / c
        dnm = acc.sum_delta_p_norm[li] / c
        dn2 = acc.sum_delta_p_norm2[li] / c
        cen = max(0.0, d2 - dm * dm)
        cenn = max(0.0, dn2 - dnm * dnm)
        out["mean_delta_p"][li] = dm
        out["mean_delta_p2"][li] = d2
        out["mean_delta_p_norm"][li] = dnm
        out["mean_delta_p_norm2"][li] = dn2
        out["dpp_raw_per_s"][li] = d2 / (2.0 * float(tau_s[li]))
        out["dpp_centered_per_s"][li] = cen / (2.0 * float(tau_s[li]))
        out["dpp_raw_norm"][li] = dn2 / (2.0 * float(tau_norm[li]))
        out["dpp_centered_norm"][li] = cenn / (2.0 * float(tau_norm[li]))
    return out


def run_dpp_full(cfg, ctx, momenta, t_s, t_norm, t_gyroperiods=None):
    """run_dpp_full (compute_dpp_full.jl:456-541) over a momenta cache shaped
    (nsave, 3, n): a DeviceBuffer, numpy / a memory map, or the h5.H5Dataset
    "momenta" of a phase-space cache file.  Returns the "dpp" and "energy_snapshots" groups as dicts."""
    t_s = np.asarray(t_s, dtype=np.float64)
    t_norm = np.asarray(t_norm, dtype=np.float64)
    t_gp = t_gyroperiods_from_axes(t_s, t_norm) if t_gyroperiods is None else np.asarray(t_gyroperiods)
    nsteps, three, total = tuple(momenta.shape)
    if three != 3:
        raise PipelineError("momenta dataset must have shape (particles, 3, steps).")
    validate_time_axes(t_s, t_norm, t_gp, key_name="D_pp trajectory time axes", require_uniform=True)
    pidx = build_particle_indices(total, cfg)
    grid = resolve_lag_grid(cfg, t_gp)
    lag_steps = grid.lag_steps
    tau_s = np.array([t_s[l] - t_s[0] for l in lag_steps])
    tau_norm = grid.tau_norm
    i0 = int(pidx[0]) - 1
    if isinstance(momenta, _lib.DeviceBuffer):
        m0 = momenta.to_host()[0, :, i0]
    elif hasattr(momenta, "read_box"):
        m0 = momenta.read_box((0, 0, i0), (1, 3, 1)).reshape(3)
    else:
        m0 = np.asarray(momenta[0, :, i0])
    p0 = momentum_magnitude(*m0)
    if not (math.isfinite(p0) and p0 > 0.0):
        raise PipelineError("Cannot compute D_pp normalization: first selected particle has invalid initial momentum.")
    acc = new_accumulators(len(lag_steps))
    snaps = selected_energy_snapshot_indices(nsteps, int(cfg.get("n_energy_snapshots", 5)))
    energy = np.asfortranarray(np.full((len(snaps), len(pidx)), np.nan))
    chunk = min(int(cfg.get("particle_chunk_size", len(pidx))), len(pidx))
    if chunk <= 0:
        raise PipelineError("particle_chunk_size must be positive.")
    batch = gpu_batch_particles(cfg, chunk, nsteps * 12)  # GPU passes over whole chunks
    if isinstance(momenta, _lib.DeviceBuffer):
        for b0 in range(0, len(pidx), batch):
            cols = pidx[b0:b0 + batch] - 1
            process_momenta_chunk_dpp(ctx, momenta, lag_steps, p0, acc, particle_index=cols)
            fill_energy_snapshots(ctx, energy, momenta, snaps, b0 + 1, particle_index=cols)
    else:  # host source (numpy, memory map, h5.H5Dataset): streamed to the device in groups of batches
        for (block,), batches in device_particle_groups(ctx, [momenta], pidx - 1, batch, [np.float32]):
            for b0, local in batches:
                process_momenta_chunk_dpp(ctx, block, lag_steps, p0, acc, particle_index=local)
                fill_energy_snapshots(ctx, energy, block, snaps, b0 + 1, particle_index=local)
    arrays = compute_dpp_arrays(acc, tau_s, tau_norm)
    dpp = dict(p0_kg_m_per_s=p0, counts=acc.counts, sum_delta_p=acc.sum_delta_p, sum_delta_p2=acc.sum_delta_p2,
               sum_delta_p_norm=acc.sum_delta_p_norm, sum_delta_p_norm2=acc.sum_delta_p_norm2, lag_steps=lag_steps,
               tau_s=tau_s, tau_norm=tau_norm, tau_gyroperiods=grid.tau_gyroperiods, **arrays)
    idx0 = np.asarray(snaps) - 1
    es = dict(snapshot_step_index=np.asarray(snaps), snapshot_t_s=t_s[idx0], snapshot_t_norm=t_norm[idx0],
              snapshot_t_gyroperiods=t_gp[idx0], energy_GeV=energy, particle_indices=pidx)
    return dict(dpp=dpp, energy_snapshots=es, lag_grid=grid)
