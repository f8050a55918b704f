"""HDF5 caches and products in the reference's layouts (libpt_h5.so).

Writers:
  finalize_cache            finalize_phase_space_writer! (phase_space_gpu_runner.jl:423-446),
                            finalize_mu_cache_writer! (multimode_multienergy_cache_pipeline.jl:1323-1347)
  save_combined_full_h5     compute_delta_mu2_dmumu_full.jl:798-950
  save_dpp_full_h5          compute_dpp_full.jl:342-437
  append_cache_metadata     multimode_multienergy_cache_pipeline.jl:1571-1631
Readers:
  validate_phase_space_cache / validate_mu_cache   (:479-528)
  open_cache                the datasets and time axes of a cache, read lazily

Scalars keep the reference's two forms: `Float64[x]` / `Int[x]` are
length-1 vectors (np.array([x])), `Int(x)` / `Float64(x)` true scalars
(Python numbers).  The cadence datasets themselves are written chunk by chunk
by the runner's writer thread (runner.CacheWriter).
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np

from . import h5
from .time_units import REFERENCE_GYROPERIOD_NAME, t_gyroperiods_from_axes, tOmega0_to_gyroperiods, \
    validate_time_axes_result


def _f64(x):
    return np.array([float(x)])


def _int(x):
    return np.array([int(x)], dtype=np.int64)


def _julia_type_name(dtype):
    return "Float64" if np.dtype(dtype) == np.float64 else "Float32"


def _sym(v):
    return str(v).lstrip(":")


# ----------------------------------------------------------- cache files ----

def write_time_reference_metadata(f, tr):
    """write_time_reference_metadata! (time_units.jl:461-477)."""
    f["Omega0"] = _f64(tr.Omega0_reference_s_inv)
    f["B0_T"] = _f64(tr.B0_reference_T)
    f["Omega0_reference_s_inv"] = _f64(tr.Omega0_reference_s_inv)
    f["B0_reference_T"] = _f64(tr.B0_reference_T)
    f["reference_gyroperiod_s"] = _f64(tr.reference_gyroperiod_s)
    f["time_reference_name"] = REFERENCE_GYROPERIOD_NAME
    f["time_reference_mode"] = str(tr.time_reference_mode)
    f["time_reference_definition"] = str(tr.time_reference_definition)
    f["time_reference_B0_definition"] = str(tr.time_reference_definition)
    f["time_reference_energy_definition"] = str(tr.time_reference_energy_definition)
    f["time_reference_source_mode"] = str(tr.time_reference_source_mode)
    f["time_reference_source_path"] = str(tr.time_reference_source_path)
    f["time_reference_source_identity"] = str(tr.time_reference_source_identity)
    f["time_reference_field_subset"] = str(tr.time_reference_field_subset)


def write_trajectory_time_metadata(f, tt, st):
    """write_trajectory_time_metadata! (time_units.jl:479-495)."""
    f["requested_trajectory_duration_gyroperiods"] = _f64(tt.requested_trajectory_duration_gyroperiods)
    f["actual_trajectory_duration_gyroperiods"] = _f64(tt.actual_trajectory_duration_gyroperiods)
    f["actual_integration_duration_gyroperiods"] = _f64(tt.actual_trajectory_duration_gyroperiods)
    f["actual_trajectory_duration_tOmega0"] = _f64(tt.actual_trajectory_duration_tOmega0)
    f["actual_trajectory_duration_s"] = _f64(tt.actual_trajectory_duration_s)
    f["requested_integration_steps_per_gyroperiod"] = _f64(tt.requested_integration_steps_per_gyroperiod)
    f["actual_integration_steps_per_gyroperiod"] = _f64(tt.actual_integration_steps_per_gyroperiod)
    f["dt_s"] = _f64(tt.dt_s)
    f["dt_tOmega0"] = _f64(tt.dt_tOmega0)
    f["dt_gyroperiods"] = _f64(tt.dt_gyroperiods)
    f["timestep_limited_by"] = str(tt.timestep_limited_by)
    f["requested_trajectory_save_interval_gyroperiods"] = _f64(st.requested_trajectory_save_interval_gyroperiods)
    f["actual_trajectory_save_interval_gyroperiods"] = _f64(st.actual_trajectory_save_interval_gyroperiods)
    f["trajectory_time_stride"] = _int(st.trajectory_time_stride)


def write_injection_metadata(f, cfg):
    """write_injection_metadata! (phase_space_gpu_runner.jl:136-143), plus
    where the draws were made (injection_rng: "host" or "device")."""
    from .runner import injection_metadata
    m = injection_metadata(cfg)
    f["injection_mode"] = str(m["injection_mode"])
    f["injection_mu0"] = np.asarray(m["injection_mu0"], dtype=np.float64)
    f["injection_position_mode"] = str(m["injection_position_mode"])
    f["injection_position"] = np.asarray(m["injection_position"], dtype=np.float64)
    f["injection_position_unit"] = str(m["injection_position_unit"])
    f["injection_rng"] = str(m["injection_rng"])


def finalize_cache(f, cfg, s, energy_GeV, alive_fraction, kind, out_dtype):
    """The metadata that finalize_phase_space_writer! / finalize_mu_cache_writer!
    append after the cadence datasets.  s: the runner's trajectory setup;
    kind: "phase" or "mu"."""
    idx = s.save_idx - 1
    t_gp = np.asarray(s.t_gp[idx], dtype=np.float64)
    f["t_norm"] = np.asarray(s.t_norm[idx], dtype=np.float64)
    f["t_s"] = np.asarray(s.t_s[idx], dtype=np.float64)
    f["t_gyroperiods"] = t_gp
    f["alive_fraction"] = np.asarray(alive_fraction[idx], dtype=np.float64)
    f["energy_GeV"] = _f64(energy_GeV)
    f["n_particles"] = _int(cfg["n_particles"])
    write_time_reference_metadata(f, s.time_reference)
    write_trajectory_time_metadata(f, s.tt, s.st)
    f["analysis_cache_duration_gyroperiods"] = _f64(t_gp[-1] - t_gp[0])
    f["integration_step_count"] = _int(s.tt.n_integration_steps)
    f["analysis_sample_count"] = _int(len(t_gp))
    f["exact_final_state_stored"] = False
    f["trajectory_mode"] = str(cfg.get("trajectory_mode", "unknown"))
    src = str(cfg.get("trajectory_field_source_path", cfg.get("file", "")))
    f["trajectory_field_source_path"] = src
    f["trajectory_field_source_identity"] = str(cfg.get("trajectory_field_source_identity", src))
    f["boundary_mode"] = str(s.boundary)
    if kind == "phase":
        f["trajectory_output_precision"] = _julia_type_name(out_dtype)
        f["position_unit"] = "m"
        f["momentum_unit"] = "kg*m/s"
    else:
        f["cache_mode"] = "mu"
        f["cache_output_precision"] = _julia_type_name(out_dtype)
        f["mu_unit"] = "dimensionless"
    write_injection_metadata(f, cfg)


def _invalid(reason, message):
    return SimpleNamespace(valid=False, reason=reason, message=str(message))


_OK = SimpleNamespace(valid=True, reason="ok", message="ok")


def _axes(f):
    t_s = np.asarray(f["t_s"], dtype=np.float64).reshape(-1)
    t_norm = np.asarray(f["t_norm"], dtype=np.float64).reshape(-1)
    t_gp = (np.asarray(f["t_gyroperiods"], dtype=np.float64).reshape(-1) if "t_gyroperiods" in f
            else t_gyroperiods_from_axes(t_s, t_norm))
    return t_s, t_norm, t_gp


def _verify_file(path):
    if not os.path.isfile(path) or os.path.getsize(path) == 0:
        raise h5.H5Error("Missing or empty file: " + str(path))


def validate_phase_space_cache(path):
    """validate_phase_space_cache_h5 (multimode_multienergy_cache_pipeline.jl:479-503)."""
    try:
        _verify_file(path)
        with h5.H5File(path, "r") as f:
            pos, mom = f.dataset("positions"), f.dataset("momenta")
            t_s, t_norm, t_gp = _axes(f)
            alive = np.asarray(f["alive_fraction"]).reshape(-1)
            msg = "Phase-space verification failed: "
            if pos.shape != mom.shape:
                return _invalid("time_length_mismatch", msg + "positions/momenta shapes differ for " + path)
            nsteps = pos.shape[0]  # Julia size(positions, 3)
            for name, v in (("t_s", t_s), ("t_norm", t_norm), ("t_gyroperiods", t_gp)):
                if len(v) != nsteps:
                    return _invalid("time_length_mismatch", msg + name + " length mismatch for " + path)
            r = validate_time_axes_result(t_s, t_norm, t_gp, key_name="Phase-space cache time axes",
                                          require_uniform=True)
            if not r.valid:
                return r
            if len(alive) != nsteps:
                return _invalid("time_length_mismatch", msg + "alive_fraction length mismatch for " + path)
            if not nsteps > 1:
                return _invalid("time_length_mismatch", msg + "not enough saved steps in " + path)
            return _OK
    except Exception as e:  # the reference returns :legacy_metadata_missing for any read failure
        return _invalid("legacy_metadata_missing", e)


def validate_mu_cache(path):
    """validate_mu_cache_h5 (multimode_multienergy_cache_pipeline.jl:505-528)."""
    try:
        _verify_file(path)
        with h5.H5File(path, "r") as f:
            mu = f.dataset("mu")
            t_s, t_norm, t_gp = _axes(f)
            alive = np.asarray(f["alive_fraction"]).reshape(-1)
            msg = "Mu-cache verification failed: "
            nsteps = mu.shape[0]  # Julia size(mu, 2)
            for name, v in (("t_s", t_s), ("t_norm", t_norm), ("t_gyroperiods", t_gp)):
                if len(v) != nsteps:
                    return _invalid("time_length_mismatch", msg + name + " length mismatch for " + path)
            r = validate_time_axes_result(t_s, t_norm, t_gp, key_name="Mu-cache time axes", require_uniform=True)
            if not r.valid:
                return r
            if len(alive) != nsteps:
                return _invalid("time_length_mismatch", msg + "alive_fraction length mismatch for " + path)
            if not mu.shape[1] > 0:
                return _invalid("time_length_mismatch", msg + "zero particles in " + path)
            if not nsteps > 1:
                return _invalid("time_length_mismatch", msg + "not enough saved steps in " + path)
            return _OK
    except Exception as e:
        return _invalid("legacy_metadata_missing", e)


def validate_cache(path, cache_mode):
    mode = _sym(cache_mode).replace("-", "_")
    return validate_mu_cache(path) if mode == "mu" else validate_phase_space_cache(path)


def open_cache(path):
    """A cache file opened for reading: the lazy cadence datasets (numpy
    shapes (nsave, n) for mu, (nsave, 3, n) for phase space) and the time
    axes.  Close with .file.close()."""
    f = h5.H5File(path, "r")
    t_s, t_norm, t_gp = _axes(f)
    out = SimpleNamespace(file=f, path=str(path), t_s=t_s, t_norm=t_norm, t_gyroperiods=t_gp,
                          alive_fraction=np.asarray(f["alive_fraction"]).reshape(-1))
    if "mu" in f:
        out.mode, out.mu = "mu", f.dataset("mu")
    else:
        out.mode, out.positions, out.momenta = "phase_space", f.dataset("positions"), f.dataset("momenta")
    return out


# ---------------------------------------------------------- product files ----

_CACHE_KEYS_PIPELINE = (
    "energy_GeV", "dt_s", "dt_tOmega0", "dt_gyroperiods", "Omega0", "B0_T", "Omega0_reference_s_inv",
    "B0_reference_T", "reference_gyroperiod_s", "time_reference_name", "time_reference_mode",
    "time_reference_definition", "time_reference_B0_definition", "time_reference_energy_definition",
    "time_reference_source_mode", "time_reference_source_path", "time_reference_source_identity",
    "time_reference_field_subset", "requested_trajectory_duration_gyroperiods",
    "actual_trajectory_duration_gyroperiods", "actual_trajectory_duration_tOmega0", "actual_trajectory_duration_s",
    "requested_integration_steps_per_gyroperiod", "actual_integration_steps_per_gyroperiod", "timestep_limited_by",
    "requested_trajectory_save_interval_gyroperiods", "actual_trajectory_save_interval_gyroperiods", "n_particles",
    "trajectory_time_stride", "analysis_cache_duration_gyroperiods", "integration_step_count",
    "analysis_sample_count", "exact_final_state_stored", "trajectory_mode", "trajectory_field_source_path",
    "trajectory_field_source_identity", "boundary_mode", "trajectory_output_precision", "cache_output_precision",
    "position_unit", "momentum_unit", "mu_unit", "injection_mode", "injection_mu0", "injection_position_mode",
    "injection_position", "injection_position_unit")

_CACHE_KEYS_DPP = (
    "energy_GeV", "dt_s", "dt_tOmega0", "dt_gyroperiods", "Omega0", "B0_T", "Omega0_reference_s_inv",
    "B0_reference_T", "reference_gyroperiod_s", "time_reference_name", "time_reference_mode",
    "time_reference_definition", "time_reference_B0_definition", "time_reference_energy_definition",
    "time_reference_source_mode", "time_reference_source_path", "time_reference_source_identity",
    "time_reference_field_subset", "requested_trajectory_duration_gyroperiods",
    "actual_trajectory_duration_gyroperiods", "actual_integration_duration_gyroperiods",
    "analysis_cache_duration_gyroperiods", "actual_trajectory_duration_tOmega0", "actual_trajectory_duration_s",
    "requested_integration_steps_per_gyroperiod", "actual_integration_steps_per_gyroperiod", "timestep_limited_by",
    "requested_trajectory_save_interval_gyroperiods", "actual_trajectory_save_interval_gyroperiods",
    "integration_step_count", "analysis_sample_count", "exact_final_state_stored", "trajectory_mode",
    "trajectory_field_source_path", "trajectory_field_source_identity", "n_particles", "trajectory_time_stride",
    "boundary_mode", "cache_mode", "momentum_unit", "injection_mode", "injection_mu0", "injection_position_mode",
    "injection_position", "injection_position_unit")


def _copy_keys(f, cache_path, keys):
    with h5.H5File(cache_path, "r") as c:
        for k in keys:
            if k in c and k not in f:
                f[k] = c[k]


def append_cache_metadata(path, cache_mode, cache_h5):
    """append_cache_metadata! (multimode_multienergy_cache_pipeline.jl:1571-1631)."""
    mode = _sym(cache_mode)
    with h5.H5File(path, "r+") as f:
        f["cache_mode"] = mode
        f["cache_h5"] = str(cache_h5)
        f["cache_dataset"] = "mu" if mode == "mu" else "positions,momenta"
        _copy_keys(f, cache_h5, _CACHE_KEYS_PIPELINE)


def copy_cache_metadata(f, cache_h5):
    """copy_cache_metadata! (compute_dpp_full.jl:332-340)."""
    if cache_h5 is None:
        return
    _copy_keys(f, cache_h5, _CACHE_KEYS_DPP)


# Defaults of the reference's standalone configs (compute_delta_mu2_dmumu_full.jl:10-48,
# compute_dpp_full.jl:15-50) for keys a caller's cfg may leave out.
_DEFAULTS = dict(compute_precision="Float32", particle_chunk_size=128, particle_selection="block_random",
                 particle_seed=20260423, particle_block_size=128, lag_mode="uniform_samples", n_lag_samples=40,
                 lag_step_stride=1, max_lag_steps=None, min_count_per_cell=20, box_length_pc=200.0,
                 field_subset=None, n_mu_bins=24, trajectory_h5="", turbulence_h5="")


def _cfg(cfg, key):
    v = cfg.get(key)
    return _DEFAULTS.get(key) if v is None else v


def _write_lag_settings(f, cfg, lag_steps, lag_grid):
    """The lag-grid settings both product files record (dmumu:848-862, dpp:361-375)."""
    f["lag_mode"] = _sym(_cfg(cfg, "lag_mode"))
    f["lag_range_policy"] = _sym(cfg.get("lag_range_policy", "fixed"))
    f["lag_common_scope"] = _sym(cfg.get("lag_common_scope", "campaign"))
    f["lag_boundary_policy"] = _sym(cfg.get("lag_boundary_policy", "strict"))
    f["max_lag_boundary_relative_error"] = _f64(cfg.get("max_lag_boundary_relative_error", 0.0))
    f["n_lag_samples"] = int(len(lag_steps))
    f["requested_n_lag_samples"] = int(_cfg(cfg, "n_lag_samples"))
    for k in ("lag_min_gyroperiods", "lag_max_gyroperiods", "lag_stride_gyroperiods"):
        if cfg.get(k) is not None:
            f[k] = _f64(cfg[k])
    f["min_lag_steps"] = int(cfg.get("min_lag_steps", 1))
    f["lag_step_stride"] = int(_cfg(cfg, "lag_step_stride"))
    f["max_lag_steps"] = -1 if cfg.get("max_lag_steps") is None else int(cfg["max_lag_steps"])
    if lag_grid is not None:
        f["source_cache_analysis_sample_count"] = _int(
            round(lag_grid.cache_lag_max_gyroperiods / lag_grid.cache_save_interval_gyroperiods) + 1)
        f["source_cache_save_interval_gyroperiods"] = _f64(lag_grid.cache_save_interval_gyroperiods)


def _write_lag_grid(f, group, lag_grid, tau_norm, n_lags):
    """Lag-mapping arrays of a product group and the file-level lag-grid
    summary (dmumu:888-921, dpp:379-410)."""
    if lag_grid is not None:
        f[group + "/requested_tau_gyroperiods"] = np.asarray(lag_grid.requested_tau_gyroperiods, dtype=np.float64)
        f[group + "/common_requested_tau_gyroperiods"] = np.asarray(lag_grid.common_requested_tau_gyroperiods,
                                                                    dtype=np.float64)
        f[group + "/tau_gyroperiods"] = np.asarray(lag_grid.tau_gyroperiods, dtype=np.float64)
        f[group + "/lag_mapping_error_gyroperiods"] = np.asarray(lag_grid.lag_mapping_error_gyroperiods,
                                                                 dtype=np.float64)
        g = lag_grid
        f["lag_mapping_max_error_gyroperiods"] = _f64(g.max_lag_mapping_error_gyroperiods)
        f["lag_mapping_max_relative_error"] = _f64(g.max_lag_mapping_relative_error)
        f["requested_lag_count"] = _int(g.requested_lag_count)
        f["unique_lag_count"] = _int(g.unique_lag_count)
        f["duplicate_lag_mapping_count"] = _int(g.duplicate_lag_mapping_count)
        f["duplicate_lag_mapping_fraction"] = _f64(g.duplicate_lag_mapping_fraction)
        for k in ("cache_lag_min_gyroperiods", "cache_lag_max_gyroperiods", "cache_save_interval_gyroperiods",
                  "configured_lag_min_gyroperiods", "configured_lag_max_gyroperiods",
                  "common_cache_lag_min_gyroperiods", "common_cache_lag_max_gyroperiods",
                  "effective_lag_min_gyroperiods", "effective_lag_max_gyroperiods"):
            f[k] = _f64(getattr(g, k))
        f["lag_comparison_group_identity"] = str(g.lag_comparison_group_identity)
        f["preflight_job_id"] = str(g.preflight_job_id)
        f["preflight_reference_group_id"] = str(g.preflight_reference_group_id)
        f["lag_group_member_count"] = _int(g.lag_group_member_count)
        f["lag_group_member_modes"] = str(g.lag_group_member_modes)
        f["lag_group_member_energies_GeV"] = str(g.lag_group_member_energies_GeV)
        f["lag_grid_source"] = _sym(g.lag_source)
    else:
        tau_gp = np.array([tOmega0_to_gyroperiods(v) for v in tau_norm])
        f[group + "/requested_tau_gyroperiods"] = tau_gp
        f[group + "/tau_gyroperiods"] = tau_gp
        f[group + "/lag_mapping_error_gyroperiods"] = np.zeros(n_lags)


def _particle_settings(f, cfg, pidx):
    pidx = np.asarray(pidx, dtype=np.int64)
    f["particle_chunk_size"] = int(_cfg(cfg, "particle_chunk_size"))
    f["first_particle"] = int(pidx[0])
    f["last_particle"] = int(pidx[-1])
    f["n_particles_used"] = int(len(pidx))
    f["particle_selection"] = _sym(_cfg(cfg, "particle_selection"))
    f["particle_seed"] = int(_cfg(cfg, "particle_seed"))
    f["particle_block_size"] = int(_cfg(cfg, "particle_block_size"))
    f["particle_indices"] = pidx


def save_combined_full_h5(path, cfg, products, *, backend="gpu", lag_grid="from_products"):
    """save_combined_full_h5 (compute_delta_mu2_dmumu_full.jl:798-950) from
    the dict that dmumu.run_combined_full / run_combined_from_mu_cache return.
    cfg["trajectory_h5"] names the source cache (the pipeline passes the mu
    cache path there, multimode_multienergy_cache_pipeline.jl:1748)."""
    from .dmumu import dmumu_start_mode, mu_bin_abs
    if lag_grid == "from_products":
        lag_grid = products.get("lag_grid")
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    dm, df = products["dmumu"], products["delta_mu2"]
    lag_steps = np.asarray(dm["lag_steps"], dtype=np.int64)
    with h5.H5File(path, "w") as f:
        f["trajectory_h5"] = str(_cfg(cfg, "trajectory_h5"))
        f["turbulence_h5"] = str(_cfg(cfg, "turbulence_h5"))
        f["compute_backend"] = _sym(backend)
        f["compute_precision"] = _sym(_cfg(cfg, "compute_precision"))
        f["source_cache_uniform_time_axis"] = True
        f["source_cache_identity"] = str(_cfg(cfg, "trajectory_h5"))
        _particle_settings(f, cfg, products["particle_indices"])
        _write_lag_settings(f, cfg, lag_steps, lag_grid)
        f["min_count_per_cell"] = int(_cfg(cfg, "min_count_per_cell"))
        f["box_length_pc"] = float(_cfg(cfg, "box_length_pc"))
        fs = cfg.get("field_subset")
        f["field_subset"] = "nothing" if fs is None else str(fs)
        f["estimated_pair_visits"] = str(int(products["estimated_pair_visits"]))
        f["tau_average_note"] = ("Lag average over the selected tau grid; this is an effective lag-averaged "
                                 "scattering measure, not a fitted diffusive plateau.")
        start_mode = dmumu_start_mode(cfg)
        bin_abs = mu_bin_abs(cfg)
        f["dmumu_start_mode"] = start_mode
        f["mu_bin_abs"] = bool(bin_abs)
        base = "mu(0)" if start_mode == "injection" else "mu(t)"
        f["mu_bin_coordinate"] = "abs(" + base + ")" if bin_abs else base
        f["n_mu_bins"] = int(cfg["n_mu_bins"])
        start_note = ("Injection-anchored accumulation: each selected particle contributes at most one pair per "
                      "selected lag, Delta mu = mu(tau) - mu(0), binned by injected mu(0)." if start_mode == "injection"
                      else "Sliding full-pair accumulation over every selected particle and every valid start time "
                           "for each selected lag; Delta mu = mu(t + tau) - mu(t), binned by pair-start mu(t).")
        bin_note = (" D_mumu bins use the absolute value of the start mu; stored mu values and Delta mu remain "
                    "signed." if bin_abs else " D_mumu bins use the signed start mu.")
        f["estimator_note"] = start_note + bin_note + " No random pair sampling."
        f.create_group("delta_mu2")
        for col in ("lag_step", "tau_s", "tau_norm", "tau_gyroperiods", "delta_mu2_particle_mean",
                    "delta_mu2_particle_std", "delta_mu2_particle_sem", "delta_mu2_pair_mean", "n_particles_used",
                    "n_pairs_used"):
            f["delta_mu2/" + col] = np.asarray(df[col])
        if lag_grid is not None:
            for k in ("requested_tau_gyroperiods", "common_requested_tau_gyroperiods",
                      "lag_mapping_error_gyroperiods"):
                f["delta_mu2/" + k] = np.asarray(getattr(lag_grid, k), dtype=np.float64)
        f.create_group("dmumu")
        f["dmumu/mu_edges"] = np.asarray(dm["mu_edges"], dtype=np.float64)
        f["dmumu/mu_centers"] = np.asarray(dm["mu_centers"], dtype=np.float64)
        f["dmumu/lag_step"] = lag_steps
        f["dmumu/lag_steps"] = lag_steps
        _write_lag_grid(f, "dmumu", lag_grid, dm["tau_norm"], len(lag_steps))
        f["dmumu/tau_s"] = np.asarray(dm["tau_s"], dtype=np.float64)
        f["dmumu/tau_norm"] = np.asarray(dm["tau_norm"], dtype=np.float64)
        # (n_lags, n_bins) numpy = Julia (n_bins, n_lags)
        f["dmumu/count_pairs_full"] = np.asarray(dm["count_pairs_full"], dtype=np.int64)
        f["dmumu/sum_delta_mu"] = np.asarray(dm["sum_delta_mu"], dtype=np.float64)
        f["dmumu/sum_delta_mu2"] = np.asarray(dm["sum_delta_mu2"], dtype=np.float64)
        for k in ("mean_delta_mu", "mean_delta_mu2", "A_mu_norm", "D_mumu_raw_norm", "D_mumu_centered_norm",
                  "D_mumu_raw_per_s", "D_mumu_centered_per_s"):
            f["dmumu/" + k] = np.asarray(dm[k], dtype=np.float64)
        for name, k in (("D_mumu_raw_tau_average_norm", "tau_avg_raw_norm"),
                        ("D_mumu_centered_tau_average_norm", "tau_avg_centered_norm"),
                        ("D_mumu_raw_tau_average_norm_count_weighted", "tau_avg_raw_norm_count_weighted"),
                        ("D_mumu_centered_tau_average_norm_count_weighted", "tau_avg_centered_norm_count_weighted")):
            f["dmumu/" + name] = np.asarray(dm[k], dtype=np.float64)


def save_dpp_full_h5(path, cfg, products, *, lag_grid="from_products"):
    """save_dpp_full_h5 (compute_dpp_full.jl:342-437) from the dict that
    dpp.run_dpp_full returns."""
    if lag_grid == "from_products":
        lag_grid = products.get("lag_grid")
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    dp, es = products["dpp"], products["energy_snapshots"]
    lag_steps = np.asarray(dp["lag_steps"], dtype=np.int64)
    traj = str(_cfg(cfg, "trajectory_h5"))
    cache = cfg.get("cache_h5") or traj
    with h5.H5File(path, "w") as f:
        f["trajectory_h5"] = traj
        f["cache_h5"] = str(cache)
        f["cache_mode"] = "phase_space"
        f["source_cache_uniform_time_axis"] = True
        f["source_cache_identity"] = str(cache)
        _particle_settings(f, cfg, es["particle_indices"])
        _write_lag_settings(f, cfg, lag_steps, lag_grid)
        f["dpp_note"] = ("Global scalar momentum diffusion uses p = sqrt(px^2 + py^2 + pz^2) and Delta p = "
                         "p(t + tau) - p(t). Normalized D_pp uses Delta p / p0 and tau * Omega0.")
        if cache and os.path.isfile(cache):
            copy_cache_metadata(f, cache)
        f.create_group("dpp")
        f["dpp/lag_step"] = lag_steps
        f["dpp/lag_steps"] = lag_steps
        _write_lag_grid(f, "dpp", lag_grid, dp["tau_norm"], len(lag_steps))
        f["dpp/tau_s"] = np.asarray(dp["tau_s"], dtype=np.float64)
        f["dpp/tau_norm"] = np.asarray(dp["tau_norm"], dtype=np.float64)
        f["dpp/p0_kg_m_per_s"] = _f64(dp["p0_kg_m_per_s"])
        f["dpp/count_pairs_full"] = np.asarray(dp["counts"], dtype=np.int64)
        for name, k in (("sum_delta_p", "sum_delta_p"), ("sum_delta_p2", "sum_delta_p2"),
                        ("sum_delta_p_over_p0", "sum_delta_p_norm"), ("sum_delta_p_over_p0_2", "sum_delta_p_norm2"),
                        ("mean_delta_p", "mean_delta_p"), ("mean_delta_p2", "mean_delta_p2"),
                        ("mean_delta_p_over_p0", "mean_delta_p_norm"), ("mean_delta_p_over_p0_2", "mean_delta_p_norm2"),
                        ("D_pp_raw_per_s", "dpp_raw_per_s"), ("D_pp_centered_per_s", "dpp_centered_per_s"),
                        ("D_pp_raw_norm", "dpp_raw_norm"), ("D_pp_centered_norm", "dpp_centered_norm")):
            f["dpp/" + name] = np.asarray(dp[k], dtype=np.float64)
        f.create_group("energy_snapshots")
        f["energy_snapshots/snapshot_step_index"] = np.asarray(es["snapshot_step_index"], dtype=np.int64)
        f["energy_snapshots/snapshot_t_s"] = np.asarray(es["snapshot_t_s"], dtype=np.float64)
        f["energy_snapshots/snapshot_t_norm"] = np.asarray(es["snapshot_t_norm"], dtype=np.float64)
        f["energy_snapshots/snapshot_t_gyroperiods"] = np.asarray(es["snapshot_t_gyroperiods"], dtype=np.float64)
        # Julia (k, N_sel) = numpy (N_sel, k) C-ordered
        f["energy_snapshots/energy_GeV"] = np.ascontiguousarray(np.asarray(es["energy_GeV"], dtype=np.float64).T)
        f["energy_snapshots/particle_indices"] = np.asarray(es["particle_indices"], dtype=np.int64)


def read_group(path, group):
    """Every dataset of a product group as a dict (numpy shapes)."""
    out = {}
    with h5.H5File(path, "r") as f:
        for name in _GROUP_KEYS.get(group, ()):
            key = group + "/" + name
            if key in f:
                out[name] = f[key]
    return out


_GROUP_KEYS = {
    "dmumu": ("mu_edges", "mu_centers", "lag_step", "lag_steps", "requested_tau_gyroperiods",
              "common_requested_tau_gyroperiods", "tau_gyroperiods", "lag_mapping_error_gyroperiods", "tau_s",
              "tau_norm", "count_pairs_full", "sum_delta_mu", "sum_delta_mu2", "mean_delta_mu", "mean_delta_mu2",
              "A_mu_norm", "D_mumu_raw_norm", "D_mumu_centered_norm", "D_mumu_raw_per_s", "D_mumu_centered_per_s",
              "D_mumu_raw_tau_average_norm", "D_mumu_centered_tau_average_norm",
              "D_mumu_raw_tau_average_norm_count_weighted", "D_mumu_centered_tau_average_norm_count_weighted"),
    "delta_mu2": ("lag_step", "tau_s", "tau_norm", "tau_gyroperiods", "delta_mu2_particle_mean",
                  "delta_mu2_particle_std", "delta_mu2_particle_sem", "delta_mu2_pair_mean", "n_particles_used",
                  "n_pairs_used", "requested_tau_gyroperiods", "common_requested_tau_gyroperiods",
                  "lag_mapping_error_gyroperiods"),
    "dpp": ("lag_step", "lag_steps", "requested_tau_gyroperiods", "common_requested_tau_gyroperiods",
            "tau_gyroperiods", "lag_mapping_error_gyroperiods", "tau_s", "tau_norm", "p0_kg_m_per_s",
            "count_pairs_full", "sum_delta_p", "sum_delta_p2", "sum_delta_p_over_p0", "sum_delta_p_over_p0_2",
            "mean_delta_p", "mean_delta_p2", "mean_delta_p_over_p0", "mean_delta_p_over_p0_2", "D_pp_raw_per_s",
            "D_pp_centered_per_s", "D_pp_raw_norm", "D_pp_centered_norm"),
    "energy_snapshots": ("snapshot_step_index", "snapshot_t_s", "snapshot_t_norm", "snapshot_t_gyroperiods",
                         "energy_GeV", "particle_indices"),
}
