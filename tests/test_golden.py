"""Replays tests/golden/runtests_known_answers.json (the known answers of the
reference's test/runtests.jl, kept as data) through the host mirror and the
oracle."""
import json
import math
import os

import numpy as np

import oracle
from particle_transport_amd.julia_compat import colon
from particle_transport_amd.time_units import (gyroperiods_to_seconds, gyroperiods_to_tOmega0, lag_cache_summary,
                                               reference_gyroperiod_s, resolve_lag_grid, resolve_save_stride,
                                               resolve_trajectory_time_grid, validate_time_axes_result)

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "runtests_known_answers.json")))
LAG = dict(lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
           max_lag_boundary_relative_error=0.0, lag_min_gyroperiods=None, lag_max_gyroperiods=None,
           lag_stride_gyroperiods=None, min_lag_steps=1, max_lag_steps=None, lag_step_stride=1, n_lag_samples=5)


def rng(d):
    return colon(d["start"], d["step"], d["stop"])


def close(a, b):
    return np.allclose(a, b, rtol=1e-12, atol=0)


def test_conversions():
    g = G["gyroperiod_conversions"]
    assert close(reference_gyroperiod_s(g["Omega0"]), g["reference_gyroperiod_s"])
    assert close(gyroperiods_to_tOmega0(1.0), g["gyroperiods_to_tOmega0_of_1"])
    assert close(gyroperiods_to_seconds(3.0, g["Omega0"]), g["gyroperiods_to_seconds_of_3"])


def test_time_grid_and_strides():
    for key in ("save_stride", "legacy_steps"):
        g = G[key]
        r = resolve_trajectory_time_grid(g["cfg"], g["Omega0"], g["v0"], g["min_dx"])
        assert resolve_save_stride(g["cfg"], r).trajectory_time_stride == g["expected_stride"]
        if "expected_eta" in g:
            assert close(r.eta_requested, g["expected_eta"])
            assert r.timestep_limited_by == g["expected_limited_by"]
        if "expected_n_integration_steps" in g:
            assert r.n_integration_steps == g["expected_n_integration_steps"]


def test_lag_grids():
    g = G["stride_lag_mapping"]
    r = resolve_lag_grid(dict(LAG, lag_mode="stride", lag_min_gyroperiods=g["lag_min_gyroperiods"],
                              lag_max_gyroperiods=g["lag_max_gyroperiods"],
                              lag_stride_gyroperiods=g["lag_stride_gyroperiods"]), rng(g["t_gyroperiods"]))
    assert list(r.lag_steps) == g["expected_lag_steps"]
    assert close(r.tau_gyroperiods, g["expected_tau_gyroperiods"])
    g = G["nearest_boundary_snap"]
    r = resolve_lag_grid(dict(LAG, lag_boundary_policy="nearest", lag_min_gyroperiods=g["lag_min_gyroperiods"],
                              lag_max_gyroperiods=g["lag_max_gyroperiods"], n_lag_samples=g["n_lag_samples"],
                              max_lag_boundary_relative_error=g["max_lag_boundary_relative_error"]),
                         rng(g["t_gyroperiods"]))
    assert close(r.tau_gyroperiods[0], g["expected_first_tau_gyroperiods"])
    g = G["first_cache_step_range"]
    r = resolve_lag_grid(dict(LAG, lag_range_policy="first_cache_step", lag_max_gyroperiods=g["lag_max_gyroperiods"],
                              n_lag_samples=g["n_lag_samples"]), rng(g["t_gyroperiods"]))
    assert close(r.effective_lag_min_gyroperiods, g["expected_effective_lag_min"])
    g = G["common_cache_intersection"]
    a, b = rng(g["a"]), rng(g["b"])
    r = resolve_lag_grid(dict(LAG, lag_range_policy="common_cache_intersection", lag_min_gyroperiods=0.25,
                              lag_max_gyroperiods=5.0, n_lag_samples=4,
                              common_cache_lag_min_gyroperiods=max(lag_cache_summary(a).cache_min_gp,
                                                                   lag_cache_summary(b).cache_min_gp),
                              common_cache_lag_max_gyroperiods=min(lag_cache_summary(a).cache_max_gp,
                                                                   lag_cache_summary(b).cache_max_gp),
                              lag_comparison_group_identity="golden"), a)
    assert close(r.effective_lag_min_gyroperiods, g["expected_effective_lag_min"])
    assert close(r.effective_lag_max_gyroperiods, g["expected_effective_lag_max"])


def test_time_axis_reasons():
    g = G["time_axis_reasons"]
    t_gp = np.array(g["t_gyroperiods"])
    t_norm = 2.0 * math.pi * t_gp
    t_s = t_norm / 4.0
    assert validate_time_axes_result(t_s, t_norm, t_gp, require_uniform=True).reason == g["nonuniform"]
    bad = t_norm.copy()
    bad[-1] += 0.1
    assert validate_time_axes_result(t_s, bad, t_gp, require_uniform=True).reason == g["units"]
    assert validate_time_axes_result(t_s[:-1], t_norm, t_gp, require_uniform=True).reason == g["length"]


def test_mean_abs_b_oracle():
    g = G["mean_abs_b"]
    b = [np.full(g["shape"], g[k]) for k in ("Bx", "By", "Bz")]
    assert oracle.mean_abs_b(*b) == g["expected_B0"]
