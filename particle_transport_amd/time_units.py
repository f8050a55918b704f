"""Reference clock, integer cadence and lag-grid resolution (host side).

Restates src/time_units.jl of the reference function by function, with the
same names, argument meaning and error behaviour: a reference ``error(...)``
raises ``PipelineError`` here.  Config dictionaries use the reference's keys
as strings (``cfg["lag_mode"]``) and its symbol values as strings
(``"uniform_samples"``); hyphenated spellings are normalised as the reference
normalises them.  Julia's range/round semantics come from ``julia_compat``.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from .julia_compat import colon, range_int_length, range_length, round_int, unique_stable

REFERENCE_GYROPERIOD_NAME = "reference gyroperiod"
TIME_REFERENCE_MODE = "total-field-mean-magnitude"
TIME_REFERENCE_B0_DEFINITION = (
    "B0_reference_T = mean(sqrt(Bx_total^2 + By_total^2 + Bz_total^2)) after magnetic-field "
    "conversion to tesla over the loaded reference grid")
TIME_REFERENCE_ENERGY_DEFINITION = (
    "Omega0_reference_s_inv = q_e * B0_reference_T / (gamma0 * m_p), where gamma0 is computed "
    "from the configured particle kinetic energy")
EPS = np.finfo(np.float64).eps


class PipelineError(RuntimeError):
    """The reference's ErrorException (error(...))."""


def _error(msg: str):
    raise PipelineError(msg)


# time_units.jl:6-10
def reference_gyroperiod_s(Omega0) -> float:
    return 2.0 * math.pi / float(Omega0)


def gyroperiods_to_tOmega0(n_gyroperiods) -> float:
    return 2.0 * math.pi * float(n_gyroperiods)


def tOmega0_to_gyroperiods(tOmega0) -> float:
    return float(tOmega0) / (2.0 * math.pi)


def gyroperiods_to_seconds(n_gyroperiods, Omega0) -> float:
    return gyroperiods_to_tOmega0(n_gyroperiods) / float(Omega0)


def seconds_to_gyroperiods(t_s, Omega0) -> float:
    return float(t_s) * float(Omega0) / (2.0 * math.pi)


# time_units.jl:12-16
def assert_finite_positive(value, key_name: str) -> float:
    v = float(value)
    if not (math.isfinite(v) and v > 0.0):
        _error(key_name + " must be finite and positive.")
    return v


def reference_B0_T(Bx, By, Bz, ctx=None) -> float:
    """time_units.jl:18-20: mean(sqrt.(Bx.^2 .+ By.^2 .+ Bz.^2)).

    The reduction runs on the GPU in FP64 (pt_mean_abs_b); ``ctx`` is a
    ``runner.Context`` (one is created if omitted)."""
    from .runner import Context
    own = ctx is None
    c = Context() if own else ctx
    try:
        return c.mean_abs_b(Bx, By, Bz)
    finally:
        if own:
            c.close()


# time_units.jl:22-24
def reference_Omega0(B0_T, gamma0, charge_C, mass_kg) -> float:
    return float(charge_C) * float(B0_T) / (float(gamma0) * float(mass_kg))


# time_units.jl:26-42
def campaign_time_reference(B0_reference_T, Omega0_reference_s_inv, *, source_mode="total", source_path="",
                            source_identity=None, field_subset="nothing"):
    if source_identity is None:
        source_identity = source_path
    Omega0 = float(Omega0_reference_s_inv)
    B0 = float(B0_reference_T)
    return SimpleNamespace(
        B0_reference_T=B0,
        Omega0_reference_s_inv=Omega0,
        reference_gyroperiod_s=reference_gyroperiod_s(Omega0),
        time_reference_mode=TIME_REFERENCE_MODE,
        time_reference_name=REFERENCE_GYROPERIOD_NAME,
        time_reference_definition=TIME_REFERENCE_B0_DEFINITION,
        time_reference_energy_definition=TIME_REFERENCE_ENERGY_DEFINITION,
        time_reference_source_mode=str(source_mode),
        time_reference_source_path=str(source_path),
        time_reference_source_identity=str(source_identity),
        time_reference_field_subset=str(field_subset),
    )


# time_units.jl:44-48
def require_explicit_campaign_time_reference(trajectory_mode, time_reference, *, caller="trajectory runner"):
    if time_reference is not None:
        return time_reference
    mode = str(trajectory_mode).lower()
    _error("Trajectory mode '" + mode + "' requires an explicit campaign time reference derived from the "
           "total-field dataset. The " + caller + " must not derive B0_reference from the trajectory field. "
           "Use a total-field wrapper for standalone total-field runs.")


def _get(cfg, key, default=None):
    v = cfg.get(key, default)
    return v


# time_units.jl:50-57
def resolve_eta_requested(cfg):
    if cfg.get("integration_steps_per_gyroperiod") is not None:
        spg = assert_finite_positive(cfg["integration_steps_per_gyroperiod"], "integration_steps_per_gyroperiod")
        return 2.0 * math.pi / spg, spg, "gyroperiods"
    eta = assert_finite_positive(cfg["eta"], "eta")
    return eta, 2.0 * math.pi / eta, "legacy_eta"


# time_units.jl:59-65
def requested_duration_gyroperiods(cfg):
    if cfg.get("trajectory_duration_gyroperiods") is not None:
        return assert_finite_positive(cfg["trajectory_duration_gyroperiods"], "trajectory_duration_gyroperiods"), \
            "gyroperiods"
    tmax = assert_finite_positive(cfg["tOmega0_max"], "tOmega0_max")
    return tOmega0_to_gyroperiods(tmax), "legacy_tOmega0"


# time_units.jl:67-111
def resolve_trajectory_time_grid(cfg, Omega0, v0, min_dx):
    eta_requested, requested_spg, integration_source = resolve_eta_requested(cfg)
    dt_gyro_s = eta_requested / float(Omega0)
    dt_cfl_s = math.inf
    limited_by = "gyro"
    dt_s = dt_gyro_s
    if bool(cfg.get("use_cfl", False)):
        cfl = assert_finite_positive(cfg["cfl"], "cfl")
        dt_cfl_s = cfl * float(min_dx) / float(v0)
        if dt_cfl_s < dt_s:
            dt_s = dt_cfl_s
            limited_by = "cfl"
    req_gp, duration_source = requested_duration_gyroperiods(cfg)
    req_s = gyroperiods_to_seconds(req_gp, Omega0)
    n_steps = int(math.floor(req_s / dt_s)) + 1
    if not n_steps > 1:
        _error("Trajectory duration resolves to fewer than two integration samples.")
    actual_s = (n_steps - 1) * dt_s
    actual_tO = actual_s * float(Omega0)
    actual_gp = tOmega0_to_gyroperiods(actual_tO)
    actual_spg = reference_gyroperiod_s(Omega0) / dt_s
    return SimpleNamespace(
        eta_requested=eta_requested,
        integration_source=integration_source,
        requested_integration_steps_per_gyroperiod=requested_spg,
        actual_integration_steps_per_gyroperiod=actual_spg,
        dt_s=dt_s,
        dt_tOmega0=dt_s * float(Omega0),
        dt_gyroperiods=seconds_to_gyroperiods(dt_s, Omega0),
        dt_gyro_s=dt_gyro_s,
        dt_cfl_s=dt_cfl_s,
        timestep_limited_by=limited_by,
        duration_source=duration_source,
        requested_trajectory_duration_gyroperiods=req_gp,
        requested_trajectory_duration_s=req_s,
        actual_trajectory_duration_gyroperiods=actual_gp,
        actual_trajectory_duration_tOmega0=actual_tO,
        actual_trajectory_duration_s=actual_s,
        n_integration_steps=n_steps,
        rounding_policy="floor(requested_duration_s / dt_s) + 1 integration samples",
    )


# time_units.jl:113-140
def resolve_save_stride(cfg, trajectory_time):
    if cfg.get("trajectory_save_interval_gyroperiods") is not None:
        req = assert_finite_positive(cfg["trajectory_save_interval_gyroperiods"],
                                     "trajectory_save_interval_gyroperiods")
        stride = max(1, round_int(req / trajectory_time.dt_gyroperiods))
        actual = stride * trajectory_time.dt_gyroperiods
        return SimpleNamespace(
            trajectory_time_stride=stride,
            save_interval_source="gyroperiods",
            requested_trajectory_save_interval_gyroperiods=req,
            actual_trajectory_save_interval_gyroperiods=actual,
            requested_trajectory_save_interval_s=req * trajectory_time.dt_s / trajectory_time.dt_gyroperiods,
            actual_trajectory_save_interval_s=stride * trajectory_time.dt_s,
            rounding_policy="round(requested_save_interval_s / dt_s) clamped to at least one integration step",
        )
    stride = int(cfg["trajectory_time_stride"])
    if not stride >= 1:
        _error("trajectory_time_stride must be >= 1")
    actual = stride * trajectory_time.dt_gyroperiods
    return SimpleNamespace(
        trajectory_time_stride=stride,
        save_interval_source="legacy_steps",
        requested_trajectory_save_interval_gyroperiods=actual,
        actual_trajectory_save_interval_gyroperiods=actual,
        requested_trajectory_save_interval_s=stride * trajectory_time.dt_s,
        actual_trajectory_save_interval_s=stride * trajectory_time.dt_s,
        rounding_policy="legacy integer trajectory_time_stride",
    )


# phase_space_gpu_runner.jl:393-396
def sampled_step_indices(nsteps: int, stride: int) -> np.ndarray:
    if stride < 1:
        _error("trajectory_time_stride must be >= 1")
    return np.arange(1, nsteps + 1, stride, dtype=np.int64)


# time_units.jl:142-149
def t_gyroperiods_from_axes(t_s, t_norm, Omega0=None):
    if t_norm is not None:
        return np.asarray(t_norm, dtype=np.float64) / (2.0 * math.pi)
    if Omega0 is not None:
        return np.asarray(t_s, dtype=np.float64) * float(Omega0) / (2.0 * math.pi)
    _error("Cannot resolve t_gyroperiods without t_norm or Omega0.")


def _isapprox(a, b, rtol, atol):
    return np.abs(a - b) <= atol + rtol * np.maximum(np.abs(a), np.abs(b))


# time_units.jl:151-169
def validate_time_axes(t_s, t_norm, t_gyroperiods, *, key_name="time axes", require_uniform=False, rtol=1.0e-10,
                       atol=1.0e-12):
    t_s = np.asarray(t_s, dtype=np.float64)
    t_norm = np.asarray(t_norm, dtype=np.float64)
    t_gp = np.asarray(t_gyroperiods, dtype=np.float64)
    n = len(t_s)
    if len(t_norm) != n:
        _error(key_name + ": t_norm length does not match t_s.")
    if len(t_gp) != n:
        _error(key_name + ": t_gyroperiods length does not match t_s.")
    if not n > 1:
        _error(key_name + ": need at least two time samples.")
    if not np.all(np.isfinite(t_s)):
        _error(key_name + ": t_s contains non-finite values.")
    if not np.all(np.isfinite(t_norm)):
        _error(key_name + ": t_norm contains non-finite values.")
    if not np.all(np.isfinite(t_gp)):
        _error(key_name + ": t_gyroperiods contains non-finite values.")
    if not np.all(np.diff(t_s) > 0.0):
        _error(key_name + ": t_s must be strictly increasing.")
    if not np.all(np.diff(t_norm) > 0.0):
        _error(key_name + ": t_norm must be strictly increasing.")
    if not np.all(np.diff(t_gp) > 0.0):
        _error(key_name + ": t_gyroperiods must be strictly increasing.")
    if not np.all(_isapprox(t_norm / (2.0 * math.pi), t_gp, 1.0e-10, 1.0e-12)):
        _error(key_name + ": t_norm and t_gyroperiods are inconsistent.")
    if require_uniform:
        intervals = np.diff(t_gp)
        si = intervals[0]
        if not np.all(np.abs(intervals - si) <= float(atol) + float(rtol) * abs(si)):
            _error(nonuniform_time_axis_message(key_name, intervals, si))
    return n


# time_units.jl:171-176
def nonuniform_time_axis_message(key_name, intervals, expected_interval):
    intervals = np.asarray(intervals, dtype=np.float64)
    diffs = np.abs(intervals - float(expected_interval))
    bad = np.nonzero(diffs > 1.0e-12 + 1.0e-10 * abs(float(expected_interval)))[0]
    first_bad = int(bad[0]) + 1 if len(bad) else 1
    return (key_name + ": nonuniform primary time axis; fixed-index lag analysis is unsafe. Likely cause: an "
            "older cache appended an off-cadence final trajectory state. Required action: regenerate the "
            "trajectory cache with the current pipeline, then rerun transport analysis. Expected interval: "
            f"{expected_interval} reference gyroperiods. First irregular interval index: {first_bad}. Observed "
            f"interval: {intervals[first_bad - 1]} reference gyroperiods.")


# time_units.jl:178-194
def validate_time_axes_result(t_s, t_norm, t_gyroperiods, *, key_name="time axes", require_uniform=True,
                              rtol=1.0e-10, atol=1.0e-12):
    t_s = np.asarray(t_s, dtype=np.float64)
    t_norm = np.asarray(t_norm, dtype=np.float64)
    t_gp = np.asarray(t_gyroperiods, dtype=np.float64)
    n = len(t_s)
    res = lambda ok, reason, msg: SimpleNamespace(valid=ok, reason=reason, message=msg)  # noqa: E731
    if len(t_norm) != n:
        return res(False, "time_length_mismatch", key_name + ": t_norm length does not match t_s.")
    if len(t_gp) != n:
        return res(False, "time_length_mismatch", key_name + ": t_gyroperiods length does not match t_s.")
    if not n > 1:
        return res(False, "time_length_mismatch", key_name + ": need at least two time samples.")
    if not (np.all(np.isfinite(t_s)) and np.all(np.isfinite(t_norm)) and np.all(np.isfinite(t_gp))):
        return res(False, "inconsistent_time_units", key_name + ": time axes contain non-finite values.")
    if not (np.all(np.diff(t_s) > 0) and np.all(np.diff(t_norm) > 0) and np.all(np.diff(t_gp) > 0)):
        return res(False, "inconsistent_time_units", key_name + ": time axes must be strictly increasing.")
    if not np.all(_isapprox(t_norm / (2.0 * math.pi), t_gp, rtol, atol)):
        return res(False, "inconsistent_time_units", key_name + ": t_norm and t_gyroperiods are inconsistent.")
    if require_uniform:
        intervals = np.diff(t_gp)
        si = intervals[0]
        if not np.all(np.abs(intervals - si) <= float(atol) + float(rtol) * abs(si)):
            return res(False, "nonuniform_time_axis", nonuniform_time_axis_message(key_name, intervals, si))
    return res(True, "ok", "ok")


# time_units.jl:196-210
def uniform_prefix_length(t_gyroperiods, *, rtol=1.0e-10, atol=1.0e-12):
    t = np.asarray(t_gyroperiods, dtype=np.float64)
    n = len(t)
    if n <= 2:
        return n
    intervals = np.diff(t)
    expected = intervals[0]
    prefix = 1
    for iv in intervals:
        if abs(iv - expected) <= atol + rtol * abs(expected):
            prefix += 1
        else:
            break
    return prefix


def _norm_symbol(value) -> str:
    return str(value).strip().lower().replace("-", "_")


# time_units.jl:212-228
def normalize_lag_boundary_policy(value):
    p = _norm_symbol(value)
    if p not in ("strict", "nearest"):
        _error("lag_boundary_policy must be strict or nearest.")
    return p


def normalize_lag_range_policy(value):
    p = _norm_symbol(value)
    if p not in ("fixed", "first_cache_step", "common_cache_intersection"):
        _error("lag_range_policy must be fixed, first-cache-step, or common-cache-intersection.")
    return p


def normalize_lag_common_scope(value):
    s = _norm_symbol(value)
    if s not in ("campaign", "reference_group"):
        _error("lag_common_scope must be campaign or reference-group.")
    return s


def _lag_mode(cfg):
    return _norm_symbol(cfg["lag_mode"])


# time_units.jl:230-239
def lag_cache_summary(t_gyroperiods):
    t = np.asarray(t_gyroperiods, dtype=np.float64)
    n = len(t)
    if not n > 1:
        _error("Need at least two cached samples to resolve lags.")
    offsets = t[1:] - t[0]
    cmin = float(offsets[0])
    cmax = float(offsets[-1])
    interval = cmin if len(offsets) == 1 else float(np.diff(t)[0])
    tol = 1.0e-12 + 1.0e-10 * max(abs(interval), abs(cmax), 1.0)
    return SimpleNamespace(offsets=offsets, cache_min_gp=cmin, cache_max_gp=cmax, cache_save_interval_gp=interval,
                           tolerance=tol)


# time_units.jl:241-248  (argmin returns the first minimiser, as np.argmin)
def nearest_positive_cache_lag(requested_gp, cache):
    req = float(requested_gp)
    if not (math.isfinite(req) and req > 0.0):
        _error("Lag boundary values must be finite and positive reference gyroperiods.")
    local = int(np.argmin(np.abs(cache.offsets - req))) + 1
    if not local >= 1:
        _error("Lag boundary cannot resolve to zero lag.")
    actual = float(cache.offsets[local - 1])
    return local, actual, abs(actual - req)


# time_units.jl:250-272
def resolve_one_lag_boundary(requested_gp, cache, cfg, boundary_name, *, analysis_label="lag"):
    policy = normalize_lag_boundary_policy(cfg.get("lag_boundary_policy", "strict"))
    req = float(requested_gp)
    step, actual, abs_err = nearest_positive_cache_lag(req, cache)
    if policy == "strict":
        if boundary_name == "minimum":
            if not req >= cache.cache_min_gp - cache.tolerance:
                _error(f"Requested lag range minimum {req} reference gyroperiods is outside representable cache "
                       f"range [{cache.cache_min_gp}, {cache.cache_max_gp}] reference gyroperiods. Use "
                       "lag_boundary_policy = \"nearest\" with an explicit tolerance for discretization-scale "
                       "cache-cadence mismatches, use lag_range_policy = \"first-cache-step\", or adjust "
                       "lag_min_gyroperiods.")
        else:
            if not req <= cache.cache_max_gp + cache.tolerance:
                _error(f"Requested lag range maximum {req} reference gyroperiods is outside representable cache "
                       f"range [{cache.cache_min_gp}, {cache.cache_max_gp}] reference gyroperiods. Use "
                       "lag_boundary_policy = \"nearest\" with an explicit tolerance for discretization-scale "
                       "cache-cadence mismatches or adjust lag_max_gyroperiods/trajectory_duration_gyroperiods.")
        return req
    max_rel = float(cfg.get("max_lag_boundary_relative_error", 0.0))
    if not (math.isfinite(max_rel) and max_rel >= 0.0):
        _error("max_lag_boundary_relative_error must be finite and nonnegative.")
    if not abs_err <= 0.5 * cache.cache_save_interval_gp + cache.tolerance:
        _error(f"Requested lag boundary {req} reference gyroperiods is too far from the nearest cached lag "
               f"{actual} reference gyroperiods for nearest policy. Absolute error {abs_err} exceeds half the "
               f"cache interval {0.5 * cache.cache_save_interval_gp}.")
    rel = abs_err / max(abs(req), EPS)
    if not rel <= max_rel + cache.tolerance:
        _error(f"Requested lag boundary {req} reference gyroperiods maps to {actual} reference gyroperiods with "
               f"relative error {rel}, exceeding max_lag_boundary_relative_error={max_rel}.")
    return actual


# time_units.jl:274-299
def configured_lag_bounds(cfg, cache):
    cmin = cfg.get("lag_min_gyroperiods")
    cmax = cfg.get("lag_max_gyroperiods")
    rp = normalize_lag_range_policy(cfg.get("lag_range_policy", "fixed"))
    if rp == "first_cache_step":
        if cmin is not None:
            _error("lag_range_policy = first-cache-step requires lag_min_gyroperiods to be absent or nothing.")
        lag_min = cache.cache_min_gp
        lag_max = cache.cache_max_gp if cmax is None else assert_finite_positive(cmax, "lag_max_gyroperiods")
        return cmin, cmax, lag_min, lag_max, math.nan, math.nan
    lag_min = cache.cache_min_gp if cmin is None else assert_finite_positive(cmin, "lag_min_gyroperiods")
    lag_max = cache.cache_max_gp if cmax is None else assert_finite_positive(cmax, "lag_max_gyroperiods")
    if rp == "common_cache_intersection":
        common_min = float(cfg.get("common_cache_lag_min_gyroperiods", math.nan))
        common_max = float(cfg.get("common_cache_lag_max_gyroperiods", math.nan))
        if not (math.isfinite(common_min) and math.isfinite(common_max)):
            _error("lag_range_policy = common-cache-intersection requires a preflight common cache range.")
        if not common_max > common_min:
            _error("Common cache lag intersection is empty.")
        lag_min = max(lag_min, common_min)
        lag_max = min(lag_max, common_max)
        if not lag_max > lag_min:
            _error("Effective lag range is empty after applying common-cache-intersection.")
        return cmin, cmax, lag_min, lag_max, common_min, common_max
    return cmin, cmax, lag_min, lag_max, math.nan, math.nan


# time_units.jl:301-315 (plus a log-spaced mode for lag grids such as 64
# log-spaced lags, which maps through the same nearest/dedupe/sort steps below)
def build_requested_lags_from_bounds(cfg, lag_min, lag_max):
    mode = _lag_mode(cfg)
    lo, hi = float(lag_min), float(lag_max)
    if not hi >= lo:
        _error("lag_max_gyroperiods must be >= lag_min_gyroperiods after policy resolution.")
    if mode == "uniform_samples":
        n = int(cfg["n_lag_samples"])
        if not n > 0:
            _error("n_lag_samples must be positive.")
        return range_length(lo, hi, n), "gyroperiods"
    if mode == "stride":
        s = assert_finite_positive(cfg["lag_stride_gyroperiods"], "lag_stride_gyroperiods")
        return colon(lo, s, hi + s * 1.0e-12), "gyroperiods"
    if mode == "log_samples":
        n = int(cfg["n_lag_samples"])
        if not n > 0:
            _error("n_lag_samples must be positive.")
        return list(np.geomspace(lo, hi, n)), "gyroperiods"
    _error(f"Unknown lag_mode: {cfg['lag_mode']}")


# time_units.jl:317-350
def requested_lag_grid_gyroperiods(cfg, max_cache_lag_step, t_gyroperiods):
    mode = _lag_mode(cfg)
    t = np.asarray(t_gyroperiods, dtype=np.float64)
    if cfg.get("lag_min_gyroperiods") is not None:
        lag_min = assert_finite_positive(cfg["lag_min_gyroperiods"], "lag_min_gyroperiods")
        if cfg.get("lag_max_gyroperiods") is not None:
            lag_max = assert_finite_positive(cfg["lag_max_gyroperiods"], "lag_max_gyroperiods")
        else:
            lag_max = float(t[max_cache_lag_step] - t[0])
        if not lag_max >= lag_min:
            _error("lag_max_gyroperiods must be >= lag_min_gyroperiods.")
        if mode in ("uniform_samples", "stride", "log_samples"):
            return build_requested_lags_from_bounds(cfg, lag_min, lag_max)
        _error(f"Unknown lag_mode: {cfg['lag_mode']}")
    min_step = max(1, int(cfg.get("min_lag_steps", 1)))
    max_step = max_cache_lag_step if cfg.get("max_lag_steps") is None else min(int(cfg["max_lag_steps"]),
                                                                                max_cache_lag_step)
    if not max_step >= min_step:
        _error("No lag steps selected. Check min_lag_steps, max_lag_steps, and saved trajectory length.")
    if mode == "uniform_samples":
        n = min(int(cfg["n_lag_samples"]), max_step - min_step + 1)
        steps = unique_stable([round_int(v) for v in range_int_length(min_step, max_step, n)])
    elif mode == "stride":
        stride = int(cfg["lag_step_stride"])
        if not stride >= 1:
            _error("lag_step_stride must be >= 1")
        steps = list(range(min_step, max_step + 1, stride))
    elif mode == "log_samples":
        n = min(int(cfg["n_lag_samples"]), max_step - min_step + 1)
        steps = unique_stable([round_int(v) for v in np.geomspace(min_step, max_step, n)])
    else:
        _error(f"Unknown lag_mode: {cfg['lag_mode']}")
    return [float(t[s] - t[0]) for s in steps], "legacy_steps"


# time_units.jl:352-459
def resolve_lag_grid(cfg, t_gyroperiods, *, min_unique_lags=2):
    t = np.asarray(t_gyroperiods, dtype=np.float64)
    nsteps = len(t)
    if not nsteps > 1:
        _error("Need at least two cached samples to resolve lags.")
    max_cache_lag_step = nsteps - 1
    cache = lag_cache_summary(t)
    rp = normalize_lag_range_policy(cfg.get("lag_range_policy", "fixed"))
    bp = normalize_lag_boundary_policy(cfg.get("lag_boundary_policy", "strict"))
    brel = float(cfg.get("max_lag_boundary_relative_error", 0.0))
    if not (math.isfinite(brel) and brel >= 0.0):
        _error("max_lag_boundary_relative_error must be finite and nonnegative.")
    label = str(cfg.get("analysis_label", "transport"))
    cmin_gp = cfg.get("lag_min_gyroperiods")
    cmax_gp = cfg.get("lag_max_gyroperiods")
    common_min = common_max = eff_min = eff_max = math.nan
    if rp == "fixed" and cmin_gp is None:
        req, source = requested_lag_grid_gyroperiods(cfg, max_cache_lag_step, t)
    else:
        cmin_gp, cmax_gp, rmin, rmax, common_min, common_max = configured_lag_bounds(cfg, cache)
        if rp == "common_cache_intersection":
            eff_min, eff_max = rmin, rmax
        else:
            eff_min = resolve_one_lag_boundary(rmin, cache, cfg, "minimum", analysis_label=label)
            eff_max = resolve_one_lag_boundary(rmax, cache, cfg, "maximum", analysis_label=label)
        if not eff_max > eff_min:
            _error("Effective lag range resolves to zero width after applying lag boundary policy.")
        req, source = build_requested_lags_from_bounds(cfg, eff_min, eff_max)
    if len(req) == 0:
        _error("No requested lag values were generated.")
    offsets = cache.offsets
    req_min = min(float(v) for v in req)
    req_max = max(float(v) for v in req)
    tol = cache.tolerance
    if bp == "strict" and (req_min < cache.cache_min_gp - tol or req_max > cache.cache_max_gp + tol):
        _error(f"Requested lag range [{req_min}, {req_max}] reference gyroperiods is outside representable cache "
               f"range [{cache.cache_min_gp}, {cache.cache_max_gp}] reference gyroperiods. Cache save interval is "
               f"{cache.cache_save_interval_gp} reference gyroperiods and cache duration is {cache.cache_max_gp} "
               "reference gyroperiods. Adjust lag_min_gyroperiods/lag_max_gyroperiods or regenerate the cache "
               "with a different trajectory_save_interval_gyroperiods or trajectory_duration_gyroperiods.")
    lag_steps, req_unique, actual_gp, errors = [], [], [], []
    seen = set()
    for r in req:
        r = float(r)
        if not (math.isfinite(r) and r > 0.0):
            _error("Requested lag values must be finite and positive reference gyroperiods.")
        local = int(np.argmin(np.abs(offsets - r)))
        step = local + 1
        if not 1 <= step <= max_cache_lag_step:
            _error("Resolved lag step is outside the cached trajectory.")
        err = abs(float(offsets[local]) - r)
        if not err <= 0.5 * cache.cache_save_interval_gp + tol:
            _error(f"Internal lag mapping inconsistency: requested {r} reference gyroperiods mapped to "
                   f"{offsets[local]}, exceeding half the cache interval {cache.cache_save_interval_gp / 2}.")
        if step not in seen:
            seen.add(step)
            lag_steps.append(step)
            req_unique.append(r)
            actual_gp.append(float(t[step] - t[0]))
            errors.append(err)
    if not len(lag_steps) >= min_unique_lags:
        _error(f"Lag grid resolved to fewer than {min_unique_lags} unique positive cached-step offsets.")
    order = np.argsort(np.asarray(lag_steps), kind="stable")
    lag_steps = np.asarray(lag_steps, dtype=np.int64)[order]
    req_unique = np.asarray(req_unique)[order]
    actual_gp = np.asarray(actual_gp)[order]
    errors = np.asarray(errors)[order]
    tau_norm = np.array([gyroperiods_to_tOmega0(v) for v in actual_gp])
    nreq = len(req)
    return SimpleNamespace(
        lag_steps=lag_steps,
        requested_tau_gyroperiods=req_unique,
        common_requested_tau_gyroperiods=np.asarray(req, dtype=np.float64),
        tau_gyroperiods=actual_gp,
        tau_norm=tau_norm,
        lag_mapping_error_gyroperiods=errors,
        max_lag_mapping_error_gyroperiods=float(errors.max()) if len(errors) else 0.0,
        max_lag_mapping_relative_error=float((errors / np.maximum(np.abs(req_unique), EPS)).max()) if len(errors)
        else 0.0,
        requested_lag_count=nreq,
        unique_lag_count=len(lag_steps),
        duplicate_lag_mapping_count=nreq - len(lag_steps),
        duplicate_lag_mapping_fraction=0.0 if nreq == 0 else (nreq - len(lag_steps)) / nreq,
        cache_lag_min_gyroperiods=cache.cache_min_gp,
        cache_lag_max_gyroperiods=cache.cache_max_gp,
        cache_save_interval_gyroperiods=cache.cache_save_interval_gp,
        lag_range_policy=rp,
        lag_boundary_policy=bp,
        max_lag_boundary_relative_error=brel,
        configured_lag_min_gyroperiods=math.nan if cmin_gp is None else float(cmin_gp),
        configured_lag_max_gyroperiods=math.nan if cmax_gp is None else float(cmax_gp),
        common_cache_lag_min_gyroperiods=common_min,
        common_cache_lag_max_gyroperiods=common_max,
        effective_lag_min_gyroperiods=eff_min if math.isfinite(eff_min) else req_min,
        effective_lag_max_gyroperiods=eff_max if math.isfinite(eff_max) else req_max,
        lag_comparison_group_identity=str(cfg.get("lag_comparison_group_identity", "not-applicable")),
        lag_common_scope=normalize_lag_common_scope(cfg.get("lag_common_scope", "campaign")),
        preflight_job_id=str(cfg.get("preflight_job_id", "not-applicable")),
        preflight_reference_group_id=str(cfg.get("preflight_reference_group_id", "not-applicable")),
        lag_group_member_count=int(cfg.get("lag_group_member_count", 1)),
        lag_group_member_modes=str(cfg.get("lag_group_member_modes", "not-applicable")),
        lag_group_member_energies_GeV=str(cfg.get("lag_group_member_energies_GeV", "not-applicable")),
        lag_source=source,
    )
