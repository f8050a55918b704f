"""Known answers of the reference's own test suite (test/runtests.jl:6-259),
ported one testset per test, against particle_transport_amd.time_units."""
import math

import numpy as np
import pytest

from particle_transport_amd.julia_compat import colon
from particle_transport_amd.time_units import (PipelineError, campaign_time_reference, gyroperiods_to_seconds,
                                               gyroperiods_to_tOmega0, lag_cache_summary,
                                               normalize_lag_common_scope, reference_gyroperiod_s,
                                               require_explicit_campaign_time_reference, resolve_lag_grid,
                                               resolve_save_stride, resolve_trajectory_time_grid,
                                               seconds_to_gyroperiods, t_gyroperiods_from_axes,
                                               tOmega0_to_gyroperiods, validate_time_axes,
                                               validate_time_axes_result)

approx = lambda a, b: np.allclose(a, b, rtol=math.sqrt(np.finfo(float).eps), atol=0)  # noqa: E731  Julia's default isapprox


def test_reference_gyroperiod_conversions():  # runtests.jl:6-13
    O = 4.0
    assert approx(reference_gyroperiod_s(O), math.pi / 2.0)
    assert approx(gyroperiods_to_tOmega0(1.0), 2.0 * math.pi)
    assert approx(tOmega0_to_gyroperiods(2.0 * math.pi), 1.0)
    assert approx(gyroperiods_to_seconds(3.0, O), 3.0 * 2.0 * math.pi / O)
    assert approx(seconds_to_gyroperiods(gyroperiods_to_seconds(2.5, O), O), 2.5)


def test_trajectory_time_resolution():  # runtests.jl:15-42
    cfg = dict(integration_steps_per_gyroperiod=64.0, eta=0.1, trajectory_duration_gyroperiods=10.0,
               tOmega0_max=100.0, use_cfl=False, cfl=0.2, trajectory_save_interval_gyroperiods=0.25,
               trajectory_time_stride=1)
    r = resolve_trajectory_time_grid(cfg, 2.0, 1.0, 1.0)
    assert approx(r.eta_requested, 2.0 * math.pi / 64.0)
    assert approx(r.requested_integration_steps_per_gyroperiod, 64.0)
    assert approx(r.actual_integration_steps_per_gyroperiod, 64.0)
    assert r.timestep_limited_by == "gyro"
    assert r.actual_trajectory_duration_gyroperiods <= 10.0
    st = resolve_save_stride(cfg, r)
    assert st.trajectory_time_stride == 16
    assert approx(st.actual_trajectory_save_interval_gyroperiods, 0.25)
    cfl = dict(cfg, use_cfl=True)
    rc = resolve_trajectory_time_grid(cfl, 2.0, 100.0, 1.0)
    assert rc.timestep_limited_by == "cfl"
    assert rc.actual_integration_steps_per_gyroperiod > rc.requested_integration_steps_per_gyroperiod


def test_legacy_equivalence():  # runtests.jl:44-61
    cfg = dict(integration_steps_per_gyroperiod=None, eta=0.1, trajectory_duration_gyroperiods=None,
               tOmega0_max=20.0, use_cfl=False, cfl=0.2, trajectory_save_interval_gyroperiods=None,
               trajectory_time_stride=3)
    r = resolve_trajectory_time_grid(cfg, 5.0, 1.0, 1.0)
    assert approx(r.eta_requested, 0.1)
    assert approx(r.requested_trajectory_duration_gyroperiods, 20.0 / (2.0 * math.pi))
    assert r.n_integration_steps == int(math.floor((20.0 / 5.0) / (0.1 / 5.0))) + 1
    assert resolve_save_stride(cfg, r).trajectory_time_stride == 3


LAG_CFG = dict(lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
               max_lag_boundary_relative_error=0.0, lag_min_gyroperiods=0.25, lag_max_gyroperiods=1.1,
               lag_stride_gyroperiods=None, min_lag_steps=1, max_lag_steps=None, lag_step_stride=1, n_lag_samples=5)


def test_lag_mapping():  # runtests.jl:63-113
    t_gp = colon(0.0, 0.25, 5.0)
    g = resolve_lag_grid(dict(LAG_CFG), t_gp)
    assert np.all(g.lag_steps >= 1)
    assert np.all(np.diff(g.lag_steps) > 0)
    assert len(np.unique(g.lag_steps)) == len(g.lag_steps)
    assert np.all(g.tau_gyroperiods > 0)
    assert approx(g.tau_norm, 2.0 * math.pi * g.tau_gyroperiods)
    s = dict(LAG_CFG, lag_mode="stride", lag_min_gyroperiods=0.25, lag_max_gyroperiods=1.0,
             lag_stride_gyroperiods=0.25)
    sg = resolve_lag_grid(s, t_gp)
    assert list(sg.lag_steps) == [1, 2, 3, 4]
    assert approx(sg.tau_gyroperiods, [0.25, 0.5, 0.75, 1.0])
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(LAG_CFG, lag_min_gyroperiods=0.01, lag_max_gyroperiods=0.5), t_gp)
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(LAG_CFG, lag_min_gyroperiods=0.5, lag_max_gyroperiods=8.0), t_gp)
    dg = resolve_lag_grid(dict(LAG_CFG, lag_min_gyroperiods=0.26, lag_max_gyroperiods=0.49, n_lag_samples=4), t_gp,
                          min_unique_lags=1)
    assert dg.requested_lag_count == 4
    assert dg.unique_lag_count < dg.requested_lag_count
    assert dg.duplicate_lag_mapping_count > 0
    assert dg.duplicate_lag_mapping_fraction > 0.0


def test_lag_boundary_policies():  # runtests.jl:115-157
    t_gp = colon(0.0, 0.016087419809019843, 5.984520168955382)
    cfg = dict(LAG_CFG, max_lag_boundary_relative_error=0.05, lag_min_gyroperiods=0.015625, lag_max_gyroperiods=5.0,
               n_lag_samples=40)
    with pytest.raises(PipelineError):
        resolve_lag_grid(cfg, t_gp)
    near = dict(cfg, lag_boundary_policy="nearest")
    g = resolve_lag_grid(near, t_gp)
    assert approx(g.tau_gyroperiods[0], 0.016087419809019843)
    assert approx(g.requested_tau_gyroperiods[0], 0.016087419809019843)
    assert approx(g.effective_lag_min_gyroperiods, 0.016087419809019843)
    assert g.lag_boundary_policy == "nearest"
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(near, max_lag_boundary_relative_error=0.01), t_gp)
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(near, lag_min_gyroperiods=0.001), t_gp)
    mx = dict(near, lag_min_gyroperiods=0.016087419809019843, lag_max_gyroperiods=t_gp[-1] + 0.25 * 0.016087419809019843)
    assert approx(resolve_lag_grid(mx, t_gp).tau_gyroperiods[-1], t_gp[-1])
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(mx, lag_max_gyroperiods=t_gp[-1] + 2.0 * 0.016087419809019843), t_gp)


def test_lag_range_policies():  # runtests.jl:159-208
    a = colon(0.0, 0.25, 5.0)
    b = colon(0.0, 0.30, 4.8)
    cfg = dict(LAG_CFG, lag_range_policy="first_cache_step", lag_min_gyroperiods=None, lag_max_gyroperiods=1.0,
               n_lag_samples=4)
    assert approx(resolve_lag_grid(cfg, b).effective_lag_min_gyroperiods, 0.30)
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(cfg, lag_min_gyroperiods=0.25), b)
    common = dict(cfg, lag_range_policy="common_cache_intersection", lag_min_gyroperiods=0.25, lag_max_gyroperiods=5.0,
                  common_cache_lag_min_gyroperiods=max(lag_cache_summary(a).cache_min_gp, lag_cache_summary(b).cache_min_gp),
                  common_cache_lag_max_gyroperiods=min(lag_cache_summary(a).cache_max_gp, lag_cache_summary(b).cache_max_gp),
                  lag_comparison_group_identity="test-group")
    cg = resolve_lag_grid(common, a)
    assert approx(cg.effective_lag_min_gyroperiods, 0.30)
    assert approx(cg.effective_lag_max_gyroperiods, 4.8)
    assert cg.lag_comparison_group_identity == "test-group"
    assert approx(cg.common_requested_tau_gyroperiods[0], 0.30)
    assert normalize_lag_common_scope("reference-group") == "reference_group"
    assert normalize_lag_common_scope("campaign") == "campaign"
    with pytest.raises(PipelineError):
        normalize_lag_common_scope("mode")
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(common, common_cache_lag_min_gyroperiods=5.0, common_cache_lag_max_gyroperiods=4.0), a)
    with pytest.raises(PipelineError):
        resolve_lag_grid(dict(common, n_lag_samples=20, lag_min_gyroperiods=0.31, lag_max_gyroperiods=0.37), a)


def test_time_axes():  # runtests.jl:210-219
    t_s = np.array([0.0, 0.5, 1.0])
    t_norm = t_s * 4.0
    t_gp = t_gyroperiods_from_axes(t_s, t_norm)
    assert validate_time_axes(t_s, t_norm, t_gp) == 3
    assert validate_time_axes(t_s, t_norm, t_gp, require_uniform=True) == 3
    with pytest.raises(PipelineError):
        validate_time_axes([0.0, 0.5, 0.4], t_norm, t_gp)
    with pytest.raises(PipelineError):
        validate_time_axes([0.0, 0.5, 1.2], [0.0, 2.0, 4.8], [0.0, 2.0 / (2 * math.pi), 4.8 / (2 * math.pi)],
                           require_uniform=True)


def test_shared_reference_metadata():  # runtests.jl:221-232 (B0 part is a GPU reduction: test_gpu_parity)
    ref = campaign_time_reference(5.0, 2.0, source_mode="total", source_path="total.h5", source_identity="total.h5",
                                  field_subset=(2, 2, 2))
    assert approx(ref.B0_reference_T, 5.0)
    assert approx(ref.Omega0_reference_s_inv, 2.0)
    assert ref.time_reference_source_mode == "total"
    assert ref.time_reference_source_path == "total.h5"
    assert approx(ref.reference_gyroperiod_s, math.pi)


def test_direct_runner_reference_guard():  # runtests.jl:234-240
    ref = campaign_time_reference(5.0, 2.0, source_mode="total", source_path="total.h5")
    assert require_explicit_campaign_time_reference("total", ref, caller="test") is ref
    for mode in ("alfven", "fast", "slow", "unknown", "total"):
        with pytest.raises(PipelineError):
            require_explicit_campaign_time_reference(mode, None, caller="test")


def test_structured_time_validation():  # runtests.jl:242-259
    t_gp = np.array([0.0, 0.25, 0.5, 0.75, 0.85])
    t_norm = 2.0 * math.pi * t_gp
    t_s = t_norm / 4.0
    r = validate_time_axes_result(t_s, t_norm, t_gp, require_uniform=True)
    assert not r.valid and r.reason == "nonuniform_time_axis"
    bad = t_norm.copy()
    bad[-1] += 0.1
    r = validate_time_axes_result(t_s, bad, t_gp, require_uniform=True)
    assert not r.valid and r.reason == "inconsistent_time_units"
    r = validate_time_axes_result(t_s[:-1], t_norm, t_gp, require_uniform=True)
    assert not r.valid and r.reason == "time_length_mismatch"
