"""<Delta mu^2>(tau) and D_mumu(mu, tau) host side over the GPU pair kernels.

Restates compute_delta_mu2_dmumu_full.jl (bins, particle selection, pair
accounting, Welford merge, finalisation) and compute_delta_mu2_curve.jl:387-429
(the per-lag dataframe).  The pair accumulation itself,
``process_mu_chunk_dmumu_cpu!`` (:656-704), runs on the GPU through
``pt_dmumu_accumulate``; ``process_mu_chunk_dmumu`` keeps its argument order.
Accumulator arrays use the reference's shapes: per-lag vectors, and the bin
arrays as Julia (n_bins, n_lags), i.e. numpy C arrays shaped (n_lags, n_bins).
"""
from __future__ import annotations

import ctypes
import math
from types import SimpleNamespace

import numpy as np

from . import _lib, h5
from ._lib import check, ptr
from .julia_compat import range_int_length, range_length, round_int, unique_stable
from .time_units import PipelineError, resolve_lag_grid, t_gyroperiods_from_axes, tOmega0_to_gyroperiods, \
    validate_time_axes


def mu_bin_abs(cfg) -> bool:
    return bool(cfg.get("mu_bin_abs", False))


def dmumu_start_mode(cfg) -> str:
    m = str(cfg.get("dmumu_start_mode", "sliding")).strip().lower().replace("-", "_")
    if m not in ("sliding", "injection"):
        raise PipelineError("dmumu_start_mode must be sliding or injection.")
    return m


def build_selected_lag_steps(nsteps: int, cfg):
    """compute_delta_mu2_dmumu_full.jl:272-287."""
    min_lag = max(1, int(cfg.get("min_lag_steps", 1)))
    max_lag = nsteps - 1 if cfg.get("max_lag_steps") is None else min(int(cfg["max_lag_steps"]), nsteps - 1)
    if not max_lag >= min_lag:
        raise PipelineError("No lag steps selected. Check min_lag_steps, max_lag_steps, and saved trajectory length.")
    mode = str(cfg["lag_mode"]).replace("-", "_")
    if mode == "uniform_samples":
        n = min(int(cfg["n_lag_samples"]), max_lag - min_lag + 1)
        return unique_stable([round_int(v) for v in range_int_length(min_lag, max_lag, n)])
    if mode == "stride":
        s = int(cfg["lag_step_stride"])
        if not s >= 1:
            raise PipelineError("lag_step_stride must be >= 1")
        return list(range(min_lag, max_lag + 1, s))
    raise PipelineError(f"Unknown lag_mode: {cfg['lag_mode']}")


def build_mu_edges_full(cfg):
    """compute_delta_mu2_dmumu_full.jl:289-302."""
    lo, hi, nb = float(cfg["mu_min"]), float(cfg["mu_max"]), int(cfg["n_mu_bins"])
    if not lo < hi:
        raise PipelineError("mu_min must be less than mu_max.")
    if not nb > 0:
        raise PipelineError("n_mu_bins must be positive.")
    if mu_bin_abs(cfg):
        if not lo >= 0.0:
            raise PipelineError("mu_min must be >= 0 when mu_bin_abs is true.")
        if not hi <= 1.0:
            raise PipelineError("mu_max must be <= 1 when mu_bin_abs is true.")
    edges = np.asarray(range_length(lo, hi, nb + 1), dtype=np.float64)
    centers = 0.5 * (edges[:-1] + edges[1:])
    return edges, centers


def mu_bin_value(mu_value, bin_abs: bool):
    return abs(mu_value) if bin_abs else mu_value


def mu_bin_index_full(mu_value, mu_edges) -> int:
    """compute_delta_mu2_dmumu_full.jl:308-318 (1-based bin, 0 = dropped)."""
    v = float(mu_value)
    if not math.isfinite(v):
        return 0
    if v < mu_edges[0] or v > mu_edges[-1]:
        return 0
    idx = int(np.searchsorted(mu_edges, v, side="right"))  # searchsortedlast, 1-based
    if idx == len(mu_edges):
        return len(mu_edges) - 1
    return 0 if idx < 1 else idx


def estimate_pair_visits(nsteps: int, selected_particle_count: int, lag_steps, start_mode="sliding") -> int:
    """compute_delta_mu2_dmumu_full.jl:331-338."""
    per = len(lag_steps) if start_mode == "injection" else sum(nsteps - int(l) for l in lag_steps)
    return int(selected_particle_count) * int(per)


def guard_exact_pair_cost(nsteps, selected_particle_count, lag_steps, cfg) -> int:
    """compute_delta_mu2_dmumu_full.jl:340-359."""
    est = estimate_pair_visits(nsteps, selected_particle_count, lag_steps, dmumu_start_mode(cfg))
    limit = int(round(float(cfg.get("max_pair_visits_without_allow", 5.0e9))))
    if est > limit and not bool(cfg.get("allow_huge", False)):
        raise PipelineError(f"Refusing to start a very large D_mumu calculation.\nEstimated pair visits: {est}\n"
                            f"Safety limit without --allow-huge: {limit}\n\nRe-run with --allow-huge if this is "
                            "intentional.")
    return est


def build_particle_indices(total_particles: int, cfg):
    """compute_delta_mu2_dmumu_full.jl:361-401, returned 1-based like the reference.

    "range" is exact.  "random" and "block-random" follow the reference's
    structure (sorted permutation prefix; sorted random blocks, truncated) with
    numpy's PCG64 in place of Julia's MersenneTwister, so their index sets differ
    from a Julia run's."""
    first = int(cfg["first_particle"])
    if not first >= 1:
        raise PipelineError("first_particle must be >= 1.")
    if not first <= total_particles:
        raise PipelineError("first_particle exceeds total particle count.")
    available = total_particles - first + 1
    req = available if cfg.get("n_particles_to_use") is None else int(cfg["n_particles_to_use"])
    if not req > 0:
        raise PipelineError("n-particles must be positive or all.")
    sel = min(available, req)
    mode = str(cfg["particle_selection"]).replace("-", "_")
    if mode == "range":
        return np.arange(first, first + sel, dtype=np.int64)
    rng = np.random.default_rng(int(cfg["particle_seed"]))
    if mode == "random":
        offsets = np.sort(rng.permutation(available)[:sel] + 1)
        return (first + offsets - 1).astype(np.int64)
    if mode == "block_random":
        bs = int(cfg["particle_block_size"])
        if not bs > 0:
            raise PipelineError("particle_block_size must be positive.")
        starts = np.arange(first, total_particles + 1, bs)
        nblk = min(len(starts), -(-sel // bs))
        chosen = np.sort(rng.permutation(len(starts))[:nblk] + 1)
        out = []
        for b in chosen:
            s0 = int(starts[b - 1])
            stop = min(total_particles, s0 + bs - 1)
            for pi in range(s0, stop + 1):
                if len(out) >= sel:
                    break
                out.append(pi)
            if len(out) >= sel:
                break
        return np.asarray(out, dtype=np.int64)
    raise PipelineError(f"Unknown particle_selection: {cfg['particle_selection']}")


def new_accumulators(n_lags: int, n_bins: int):
    """The zeroed accumulators of run_combined_full / run_combined_from_mu_cache
    (multimode_multienergy_cache_pipeline.jl:1674-1681)."""
    return SimpleNamespace(
        particle_counts=np.zeros(n_lags, np.int64), particle_means=np.zeros(n_lags),
        particle_m2s=np.zeros(n_lags), pair_sum_squares=np.zeros(n_lags), pair_counts=np.zeros(n_lags, np.int64),
        dmumu_counts=np.zeros((n_lags, n_bins), np.int64), dmumu_sum_delta=np.zeros((n_lags, n_bins)),
        dmumu_sum_delta2=np.zeros((n_lags, n_bins)))


def process_mu_chunk_dmumu(ctx, mu, lag_steps, mu_edges, acc, start_mode="sliding", bin_abs=False, *,
                           layout="time_fastest", particle_index=None, n_sel=None):
    """process_mu_chunk_dmumu_cpu! (compute_delta_mu2_dmumu_full.jl:656-704) on the GPU.

    mu: float32 numpy array, or a _lib.DeviceBuffer already in HBM.
      layout "time_fastest": the reference's mu_chunk, numpy shape (chunk, nsave)
        (Julia (nsave, chunk)); column j is particle j.
      layout "particle_fastest": a mu cache (nsave, n) as the pushers write it;
        particle_index (0-based) selects columns, else the first n_sel.
    acc: from new_accumulators, updated in place."""
    if isinstance(mu, _lib.DeviceBuffer):
        if mu.dtype != np.float32:
            raise PipelineError("mu must be float32")
        on_dev = 1
        rows, cols = mu.shape
        p = ctypes.cast(ctypes.c_void_p(mu.ptr), _lib._fp)
    else:
        on_dev = 0
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        rows, cols = mu.shape
        p = mu.ctypes.data_as(_lib._fp)
    if layout == "time_fastest":
        chunk, nsave = rows, cols
        ss, sp = 1, nsave
        total = chunk
    else:
        nsave, total = rows, cols
        ss, sp = total, 1
    idx = None if particle_index is None else np.ascontiguousarray(particle_index, dtype=np.int64)
    if idx is not None:
        if len(idx) and (idx.min() < 0 or idx.max() >= total):
            raise PipelineError("particle index out of range")
        nsel = len(idx)
    else:
        nsel = total if n_sel is None else int(n_sel)
        if nsel > total:
            raise PipelineError("n_sel exceeds the particle count")
    lags = np.ascontiguousarray(lag_steps, dtype=np.int64)
    edges = np.ascontiguousarray(mu_edges, dtype=np.float64)
    mode = {"sliding": 0, "injection": 1}[str(start_mode).replace("-", "_")]
    check(_lib.load().pt_dmumu_accumulate(
        ctx.h, p, on_dev, nsave, nsel, ss, sp, ptr(idx, _lib._i64p) if idx is not None else None,
        ptr(lags, _lib._i64p), len(lags), ptr(edges, _lib._dp), len(edges) - 1, mode, int(bool(bin_abs)),
        ptr(acc.particle_counts, _lib._i64p), ptr(acc.particle_means, _lib._dp), ptr(acc.particle_m2s, _lib._dp),
        ptr(acc.pair_sum_squares, _lib._dp), ptr(acc.pair_counts, _lib._i64p), ptr(acc.dmumu_counts, _lib._i64p),
        ptr(acc.dmumu_sum_delta, _lib._dp), ptr(acc.dmumu_sum_delta2, _lib._dp)))
    return acc


def compute_dmumu_arrays_full(counts, sum_delta, sum_delta2, tau_s, tau_norm, cfg):
    """compute_delta_mu2_dmumu_full.jl:706-744.  Inputs/outputs (n_lags, n_bins)."""
    min_count = int(cfg["min_count_per_cell"])
    counts = np.asarray(counts)
    shape = counts.shape
    out = [np.full(shape, np.nan) for _ in range(7)]
    mean_delta, mean_delta2, drift, raw_n, cen_n, raw_s, cen_s = out
    for li in range(shape[0]):
        tn, ts = float(tau_norm[li]), float(tau_s[li])
        for b in range(shape[1]):
            c = counts[li, b]
            if not c >= min_count:
                continue
            dm = sum_delta[li, b] / c
            d2 = sum_delta2[li, b] / c
            centered = max(0.0, d2 - dm * dm)
            mean_delta[li, b] = dm
            mean_delta2[li, b] = d2
            drift[li, b] = dm / tn
            raw_n[li, b] = d2 / (2.0 * tn)
            cen_n[li, b] = centered / (2.0 * tn)
            raw_s[li, b] = d2 / (2.0 * ts)
            cen_s[li, b] = centered / (2.0 * ts)
    return tuple(out)


def average_over_tau_full(values, counts, cfg):
    """compute_delta_mu2_dmumu_full.jl:746-776.  values/counts (n_lags, n_bins)."""
    min_count = int(cfg["min_count_per_cell"])
    nl, nb = np.shape(values)
    unweighted = np.full(nb, np.nan)
    weighted = np.full(nb, np.nan)
    for b in range(nb):
        vs, nv, ws, tw = 0.0, 0, 0.0, 0
        for li in range(nl):
            v, c = values[li, b], counts[li, b]
            if math.isfinite(v) and c >= min_count:
                vs += v
                nv += 1
                ws += v * c
                tw += c
        if nv > 0:
            unweighted[b] = vs / nv
        if tw > 0:
            weighted[b] = ws / tw
    return unweighted, weighted


def sem_from_stats(count: int, m2: float) -> float:
    """compute_delta_mu2_curve.jl:387-393."""
    if count <= 1:
        return math.nan
    return math.sqrt(m2 / (count - 1) / count)


def build_output_dataframe(lag_steps, t_s, t_norm, acc):
    """compute_delta_mu2_curve.jl:409-429, as a dict of columns."""
    t_s = np.asarray(t_s, dtype=np.float64)
    t_norm = np.asarray(t_norm, dtype=np.float64)
    lag = np.asarray(lag_steps, dtype=np.int64)
    tau_s = t_s[lag] - t_s[0]
    tau_norm = t_norm[lag] - t_norm[0]
    pc = acc.particle_counts
    std = np.array([math.nan if c <= 1 else math.sqrt(acc.particle_m2s[i] / (c - 1)) for i, c in enumerate(pc)])
    sem = np.array([sem_from_stats(int(c), acc.particle_m2s[i]) for i, c in enumerate(pc)])
    pair_mean = np.array([math.nan if acc.pair_counts[i] == 0 else acc.pair_sum_squares[i] / acc.pair_counts[i]
                          for i in range(len(pc))])
    return dict(lag_step=lag, tau_s=tau_s, tau_norm=tau_norm,
                tau_gyroperiods=np.array([tOmega0_to_gyroperiods(v) for v in tau_norm]),
                delta_mu2_particle_mean=acc.particle_means.copy(), delta_mu2_particle_std=std,
                delta_mu2_particle_sem=sem, delta_mu2_pair_mean=pair_mean, n_particles_used=pc.copy(),
                n_pairs_used=acc.pair_counts.copy())


def _finish_products(cfg, acc, lag_steps, grid, tau_s, tau_norm, t_s, t_norm, edges, centers, visits, pidx):
    """The "delta_mu2" and "dmumu" groups of delta_mu2_dmumu_full.h5 from the
    merged accumulators (compute_delta_mu2_dmumu_full.jl:1180-1196)."""
    df = build_output_dataframe(lag_steps, t_s, t_norm, acc)
    arrays = compute_dmumu_arrays_full(acc.dmumu_counts, acc.dmumu_sum_delta, acc.dmumu_sum_delta2, tau_s, tau_norm,
                                       cfg)
    names = ("mean_delta_mu", "mean_delta_mu2", "A_mu_norm", "D_mumu_raw_norm", "D_mumu_centered_norm",
             "D_mumu_raw_per_s", "D_mumu_centered_per_s")
    dm = dict(zip(names, arrays))
    dm["tau_avg_raw_norm"], dm["tau_avg_raw_norm_count_weighted"] = average_over_tau_full(arrays[3],
                                                                                          acc.dmumu_counts, cfg)
    dm["tau_avg_centered_norm"], dm["tau_avg_centered_norm_count_weighted"] = average_over_tau_full(
        arrays[4], acc.dmumu_counts, cfg)
    dm.update(mu_edges=edges, mu_centers=centers, lag_steps=lag_steps, tau_s=tau_s, tau_norm=tau_norm,
              tau_gyroperiods=grid.tau_gyroperiods, count_pairs_full=acc.dmumu_counts,
              sum_delta_mu=acc.dmumu_sum_delta, sum_delta_mu2=acc.dmumu_sum_delta2)
    return dict(delta_mu2=df, dmumu=dm, estimated_pair_visits=visits, particle_indices=pidx, lag_grid=grid,
                accumulators=acc)


def _setup_pass(cfg, t_s, t_norm, t_gyroperiods, nsteps, total, key_name):
    edges, centers = build_mu_edges_full(cfg)
    t_s = np.asarray(t_s, dtype=np.float64)
    t_norm = np.asarray(t_norm, dtype=np.float64)
    t_gp = t_gyroperiods_from_axes(t_s, t_norm) if t_gyroperiods is None else np.asarray(t_gyroperiods)
    validate_time_axes(t_s, t_norm, t_gp, key_name=key_name, require_uniform=True)
    pidx = build_particle_indices(total, cfg)
    if len(pidx) == 0:
        raise PipelineError("No particles selected.")
    grid = resolve_lag_grid(cfg, t_gp)
    lag_steps = grid.lag_steps
    tau_s = np.array([t_s[l] - t_s[0] for l in lag_steps])
    visits = guard_exact_pair_cost(nsteps, len(pidx), lag_steps, cfg)
    chunk = min(int(cfg.get("particle_chunk_size", len(pidx))), len(pidx))
    if chunk <= 0:
        raise PipelineError("particle_chunk_size must be positive.")
    return SimpleNamespace(edges=edges, centers=centers, t_s=t_s, t_norm=t_norm, pidx=pidx, grid=grid,
                           lag_steps=lag_steps, tau_s=tau_s, tau_norm=grid.tau_norm, visits=visits, chunk=chunk,
                           acc=new_accumulators(len(lag_steps), len(centers)))


def gpu_batch_particles(cfg, chunk, bytes_per_particle, budget_bytes=256 << 20):
    """Particles per GPU pass: whole analysis chunks (particle_chunk_size),
    up to cfg["gpu_batch_particles"] (default 65536) and at most budget_bytes
    of input.  The reference's chunk of 128 particles would leave most of the
    GPU idle (a 128-particle D_mumu call fills 32 workgroups); batches of
    chunks merge with Chan's formula in order, as the reference merges chunks
    and threads, so only the rounding of the Welford merge tree differs."""
    want = int(cfg.get("gpu_batch_particles", 65536))
    if want <= 0:
        raise PipelineError("gpu_batch_particles must be positive.")
    cap = max(chunk, int(budget_bytes) // max(1, int(bytes_per_particle)))
    return max(chunk, (min(max(want, chunk), cap) // chunk) * chunk)


def _direct_sources(sources, dtypes):
    """True when every source is a C-contiguous page-locked host array
    (a cache the runner keeps in host memory) of the wanted dtype: the
    device's copy engine then reads a group's columns in place."""
    return all(isinstance(s, np.ndarray) and s.flags.c_contiguous and s.dtype == np.dtype(dt) and _lib.is_pinned(s)
               for s, dt in zip(sources, dtypes))


def _direct_groups(ctx, sources, pidx0, chunk, dtypes, budget_bytes):
    """device_particle_groups over page-locked sources: a worker thread has
    the copy engine move the next group's columns into one of two sets of
    device buffers, on a context (stream) of its own, while the caller's
    kernels work on the current group.  A caller must be done with a group's
    device views when it asks for the group after next (the statistics
    passes synchronise before they return)."""
    from concurrent.futures import ThreadPoolExecutor

    from .runner import Context

    pidx0 = np.asarray(pidx0, dtype=np.int64)
    per_col = max(int(np.prod(s.shape[:-1])) * s.dtype.itemsize for s in sources)
    plans = h5.plan_particle_groups(pidx0, chunk, max(1, int(budget_bytes) // max(1, per_col)))
    cctx = Context(getattr(ctx, "device", 0))
    sets = [[None] * len(sources) for _ in range(2)]

    def dev(slot, k, shape, dt):
        n = max(1, int(np.prod(shape)))
        b = sets[slot][k]
        if b is None or b.nbytes < n * np.dtype(dt).itemsize:
            if b is not None:
                b.free()
            b = sets[slot][k] = _lib.DeviceBuffer(cctx, (n,), dt)
        return b.view(shape)

    def upload(plan, slot):
        kind, g0, g1, lo, hi = plan
        views = []
        for k, (s, dt) in enumerate(zip(sources, dtypes)):
            if kind == "span":
                v = dev(slot, k, tuple(s.shape[:-1]) + (hi - lo + 1,), dt)
                v.upload_columns(s, lo)  # synchronous on the copy context
            else:
                a = np.ascontiguousarray(np.take(s, pidx0[g0:g1], axis=-1), dtype=dt)
                v = dev(slot, k, a.shape, dt)
                v.upload(a)
            views.append(v)
        return views, h5.plan_chunks(plan, pidx0, chunk)

    fut = None
    try:
        with ThreadPoolExecutor(max_workers=1) as ex:
            fut = ex.submit(upload, plans[0], 0) if plans else None
            for k in range(len(plans)):
                cur = fut.result()
                fut = ex.submit(upload, plans[k + 1], (k + 1) & 1) if k + 1 < len(plans) else None
                yield cur
    finally:
        if fut is not None:
            fut.cancel()
            try:
                fut.result()
            except Exception:  # the copy of a group nobody asked for
                pass
        for bs in sets:
            for b in bs:
                if b is not None:
                    b.free()
        cctx.close()


def device_particle_groups(ctx, sources, pidx0, chunk, dtypes, budget_bytes=256 << 20):
    """h5.iter_particle_groups over host sources (numpy arrays, memory maps
    or h5.H5Datasets), each group uploaded once into reusable device buffers:
    yields (device views, chunks).  The next group is read from the host
    source while the GPU works on the current one.  Page-locked sources of
    the wanted dtype skip the host read: a group's span of columns goes to
    the device in one strided copy per source, the next group's while the
    GPU works on the current one (_direct_groups)."""
    bufs = [None] * len(sources)

    def dev(k, shape, dt):
        n = max(1, int(np.prod(shape)))
        if bufs[k] is None or bufs[k].nbytes < n * np.dtype(dt).itemsize:
            if bufs[k] is not None:
                bufs[k].free()
            bufs[k] = _lib.DeviceBuffer(ctx, (n,), dt)
        return bufs[k].view(shape)

    if _direct_sources(sources, dtypes):
        yield from _direct_groups(ctx, sources, pidx0, chunk, dtypes, budget_bytes)
        return

    srcs = [s if hasattr(s, "read_box") else h5.ArrayDataset(np.asarray(s)) for s in sources]
    try:
        for arrays, chunks in h5.iter_particle_groups(srcs, pidx0, chunk, budget_bytes, reuse=True):
            views = []
            for k, (a, dt) in enumerate(zip(arrays, dtypes)):
                a = np.ascontiguousarray(a, dtype=dt)
                v = dev(k, a.shape, dt)
                v.upload(a)
                views.append(v)
            yield views, chunks
    finally:
        for b in bufs:
            if b is not None:
                b.free()


def run_combined_from_mu_cache(cfg, ctx, mu_cache, t_s, t_norm, t_gyroperiods=None):
    """run_combined_from_mu_cache (multimode_multienergy_cache_pipeline.jl:1633-1812)
    over a mu cache shaped (nsave, n) (Julia (n, nsave)): a DeviceBuffer
    already on the device, or a host source (numpy, a memory map, or the
    h5.H5Dataset "mu" of a cache file) streamed to the device in groups of
    chunks, the next group read while the GPU works on the current one.
    Each GPU pass covers a batch of whole chunks (gpu_batch_particles).
    Returns the "delta_mu2" and "dmumu" product groups of
    delta_mu2_dmumu_full.h5 as dicts."""
    nsteps, total = mu_cache.shape
    u = _setup_pass(cfg, t_s, t_norm, t_gyroperiods, nsteps, total, "Mu-cache time axes")
    start_mode, bin_abs = dmumu_start_mode(cfg), mu_bin_abs(cfg)
    # Batches of whole chunks, merged with Chan's formula in order.
    batch = gpu_batch_particles(cfg, u.chunk, nsteps * 4)
    if isinstance(mu_cache, _lib.DeviceBuffer):
        for b0 in range(0, len(u.pidx), batch):
            process_mu_chunk_dmumu(ctx, mu_cache, u.lag_steps, u.edges, u.acc, start_mode, bin_abs,
                                   layout="particle_fastest", particle_index=u.pidx[b0:b0 + batch] - 1)
    else:
        for (block,), batches in device_particle_groups(ctx, [mu_cache], u.pidx - 1, batch, [np.float32]):
            for _, local in batches:
                process_mu_chunk_dmumu(ctx, block, u.lag_steps, u.edges, u.acc, start_mode, bin_abs,
                                       layout="particle_fastest", particle_index=local)
    return _finish_products(cfg, u.acc, u.lag_steps, u.grid, u.tau_s, u.tau_norm, u.t_s, u.t_norm, u.edges,
                            u.centers, u.visits, u.pidx)


def reconstruct_mu_chunk(ctx, field, positions, momenta, out=None, sample_major=False):
    """reconstruct_mu_chunk_gpu (compute_delta_mu2_curve.jl:186-225): mu for
    every (sample, particle) of a phase-space chunk, B trilinear from a planar
    field in the field's precision.  positions/momenta: numpy (nsave, 3,
    chunk) (Julia (chunk, 3, nsave)) or DeviceBuffers of that shape.  Writes
    the reference's mu_chunk, Julia (nsave, chunk) = numpy (chunk, nsave)
    float32, into `out` (a DeviceBuffer, kept on the device) or returns it as
    a host array.  sample_major=True writes numpy (nsave, chunk) instead, the
    layout the pair walk reads coalesced (process_mu_chunk_dmumu with
    layout="particle_fastest")."""
    keep = []

    def src(a):
        if isinstance(a, _lib.DeviceBuffer):
            return ctypes.c_void_p(a.ptr), (_lib.PT_F64 if a.dtype == np.float64 else _lib.PT_F32), 1, a.shape
        arr = np.ascontiguousarray(a)
        if arr.dtype not in (np.float32, np.float64):
            arr = arr.astype(np.float64)
        keep.append(arr)
        return ptr(arr), (_lib.PT_F64 if arr.dtype == np.float64 else _lib.PT_F32), 0, arr.shape

    pp, pdt, pdev, pshape = src(positions)
    mp, mdt, mdev, mshape = src(momenta)
    if pshape != mshape or len(pshape) != 3 or pshape[1] != 3:
        raise PipelineError("positions and momenta must both have shape (particles, 3, steps).")
    if (pdt, pdev) != (mdt, mdev):
        raise PipelineError("positions and momenta must share dtype and device")
    nsave, _, chunk = pshape
    shape = (nsave, chunk) if sample_major else (chunk, nsave)
    if out is None:
        host = np.empty(shape, dtype=np.float32)
        dst, on_dev = ptr(host), 0
    else:
        if out.dtype != np.float32 or out.shape != shape:
            raise PipelineError(f"out must be a float32 DeviceBuffer shaped {shape}")
        host, dst, on_dev = None, ctypes.c_void_p(out.ptr), 1
    check(_lib.load().pt_mu_from_phase_layout(ctx.h, field.h, chunk, nsave, pdt, pdev, pp, mp, _lib.PT_F32, on_dev,
                                              dst, 1 if sample_major else 0))
    return out if host is None else host


def run_combined_full(cfg, ctx, field, positions, momenta, t_s, t_norm, t_gyroperiods=None):
    """run_combined_full (compute_delta_mu2_dmumu_full.jl:1012-1259) over a
    phase-space cache: per batch of particle chunks (gpu_batch_particles), mu
    is reconstructed on the device
    (reconstruct_mu_chunk) into a device buffer that the pair pass reads
    directly, so mu never travels to the host.  positions/momenta: numpy
    (nsave, 3, n) as the phase-space cache stores them (Julia (n, 3, nsave)),
    e.g. memory maps; each chunk's columns are gathered on the host like
    read_particle_batch (:411-437), or h5.H5Datasets of a phase-space cache
    file, streamed in groups of chunks.  A page-locked cache (the runner's
    in-memory form) is not read by the host at all: the copy engine moves
    each batch's columns straight to the device.  field: a planar Field with B (the
    reference loads B in compute_precision, Float32 by default).  Returns the
    product groups like run_combined_from_mu_cache."""
    nsteps, three, total = tuple(positions.shape)
    if three != 3 or tuple(momenta.shape) != (nsteps, 3, total):
        raise PipelineError("positions and momenta must both have shape (particles, 3, steps).")
    if field.storage not in ("f32_planar", "f64_planar"):
        raise PipelineError("mu reconstruction reads a planar field (f32_planar or f64_planar)")
    u = _setup_pass(cfg, t_s, t_norm, t_gyroperiods, nsteps, total, "Trajectory HDF5 time axes")
    start_mode, bin_abs = dmumu_start_mode(cfg), mu_bin_abs(cfg)
    T = np.float64 if field.storage == "f64_planar" else np.float32
    batch = gpu_batch_particles(cfg, u.chunk, nsteps * 6 * np.dtype(T).itemsize)
    buf = _lib.DeviceBuffer(ctx, (nsteps, min(batch, len(u.pidx))), np.float32)

    def one_batch(pos, mom):
        if not isinstance(pos, _lib.DeviceBuffer):
            pos, mom = np.ascontiguousarray(pos, dtype=T), np.ascontiguousarray(mom, dtype=T)
        # sample-major on the device: the pair walk then reads it coalesced
        mu = reconstruct_mu_chunk(ctx, field, pos, mom, out=buf.view((nsteps, pos.shape[-1])), sample_major=True)
        process_mu_chunk_dmumu(ctx, mu, u.lag_steps, u.edges, u.acc, start_mode, bin_abs, layout="particle_fastest")

    def whole(local, width):
        return len(local) == width and (len(local) == 0 or (local[0] == 0 and local[-1] == len(local) - 1))

    # One batch per group: the group's arrays are then the batch itself and
    # go to the device without a second gather (a scattered selection, whose
    # group spans more columns than it selects, is still gathered on the host).
    try:
        if _direct_sources((positions, momenta), (T, T)):
            # page-locked cache: the copy engine reads each group's columns in place
            per_col = nsteps * 3 * np.dtype(T).itemsize
            for (P, M), batches in device_particle_groups(ctx, (positions, momenta), u.pidx - 1, batch, (T, T),
                                                          budget_bytes=batch * per_col):
                for c0, local in batches:
                    if whole(local, P.shape[-1]):
                        one_batch(P, M)
                    else:
                        cols = u.pidx[c0:c0 + len(local)] - 1
                        one_batch(np.take(positions, cols, axis=-1), np.take(momenta, cols, axis=-1))
        else:
            srcs = [a if hasattr(a, "read_box") else h5.ArrayDataset(np.asarray(a)) for a in (positions, momenta)]
            per_col = max(int(np.prod(s.shape[:-1])) * np.dtype(s.dtype).itemsize for s in srcs)
            for (P, M), batches in h5.iter_particle_groups(srcs, u.pidx - 1, batch, budget_bytes=batch * per_col,
                                                           reuse=True):
                for _, local in batches:
                    if whole(local, P.shape[-1]):
                        one_batch(P, M)
                    else:
                        one_batch(P[:, :, local], M[:, :, local])
    finally:
        buf.free()
    return _finish_products(cfg, u.acc, u.lag_steps, u.grid, u.tau_s, u.tau_norm, u.t_s, u.t_norm, u.edges,
                            u.centers, u.visits, u.pidx)
