"""End to end through the reference's entry points (runner / dmumu / dpp
mirrors) on the GPU, checked against the oracle on the same initial state."""
import json
import os

import numpy as np
import pytest

import oracle
from particle_transport_amd import dmumu, dpp, fields, runner
from particle_transport_amd.injection import init_particles

pytestmark = pytest.mark.gpu


def base_cfg(**kw):
    cfg = dict(box_length_pc=10.0, boundary="periodic", precision="Float64", integration_steps_per_gyroperiod=50.0,
               trajectory_duration_gyroperiods=4.0, trajectory_save_interval_gyroperiods=0.2, use_cfl=False, cfl=0.2,
               eta=0.1, tOmega0_max=1.0, trajectory_time_stride=1, seed=5, n_particles=300, trajectory_mode="total",
               injection_mode="isotropic", mu_cache_output_precision="Float32",
               trajectory_output_precision="Float32", progress_every=0)
    cfg.update(kw)
    return cfg


@pytest.fixture(scope="module")
def field17():
    raw = fields.generate_turbulence(17, seed=8, kmax=4, dB_over_B0=0.7)["total"]
    return [np.asfortranarray(a) for a in fields.to_si(raw)]


def oracle_for(cfg, F, res, energy, out_kind):
    s = runner._trajectory_setup(cfg, F, energy, res.time_reference, "test")
    x1, x2, x3, p1, p2, p3 = init_particles(cfg, s.axes, s.gamma0, s.v0, s.T, F[0], F[1], F[2])
    x = np.ascontiguousarray(np.stack([x1, x2, x3]), dtype=np.float64)
    p = np.ascontiguousarray(np.stack([p1, p2, p3]), dtype=np.float64)
    g = oracle.make_geom(17, 17, 17, fields.box_length_m(10.0))
    return s, oracle.run(F, g, x, p, float(s.dt), 1, s.nsteps, s.nsteps, s.st.trajectory_time_stride,
                         s.boundary == "periodic", out_kind)


def test_mu_cache_runner_and_dmumu_product(tmp_path, field17):
    F = field17
    # 4.01 T_g0: floor(duration / dt) is 200 well away from an integer edge, so
    # the last bit of B0 cannot move the step count
    cfg = base_cfg(trajectory_duration_gyroperiods=4.01)
    res = runner.run_total_field_mu_cache_ensemble(cfg, F, str(tmp_path / "mu"), energy_GeV=10.0, chunk_saves=3)
    s, o = oracle_for(cfg, F, res, 10.0, 2)
    mu = np.asarray(res.cache["mu"])
    assert mu.shape == (len(s.save_idx), 300) == o["mu"].shape
    assert np.max(np.abs(mu - o["mu"].astype(np.float32))) <= 1e-6  # FP64 path, float32 cache
    assert s.st.trajectory_time_stride == 10 and s.nsteps == 201  # 4.01 T_g0 at 50 steps/T_g0, save every 0.2 T_g0
    meta = json.load(open(os.path.join(str(tmp_path / "mu"), "metadata.json")))
    assert meta["cache_mode"] == "mu" and meta["integration_step_count"] == 201
    assert np.allclose(meta["alive_fraction"], 1.0)
    # D_mumu product over the cache, range selection, vs the oracle on the same mu
    dcfg = dict(mu_min=0.0, mu_max=1.0, n_mu_bins=10, mu_bin_abs=True, dmumu_start_mode="sliding",
                first_particle=1, n_particles_to_use=None, particle_selection="range", particle_chunk_size=128,
                lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
                lag_min_gyroperiods=None, lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None,
                n_lag_samples=8, min_count_per_cell=5, max_pair_visits_without_allow=5e9, allow_huge=False)
    with runner.Context(0) as ctx:
        prod = dmumu.run_combined_from_mu_cache(dcfg, ctx, mu, meta["t_s"], meta["t_norm"])
    acc = oracle.new_dmumu_accumulators(len(prod["dmumu"]["lag_steps"]), 10)
    for c0 in range(0, 300, 128):
        oracle.dmumu(mu.T[c0:c0 + 128], prod["dmumu"]["lag_steps"], prod["dmumu"]["mu_edges"], False, True, acc)
    assert np.array_equal(prod["dmumu"]["count_pairs_full"], acc["dmumu_counts"])
    assert np.allclose(prod["dmumu"]["sum_delta_mu2"], acc["dmumu_sum_delta2"], rtol=1e-11)
    assert np.allclose(prod["delta_mu2"]["delta_mu2_particle_mean"], acc["particle_means"], rtol=1e-11)
    assert np.array_equal(prod["delta_mu2"]["n_pairs_used"], acc["pair_counts"])


def test_phase_space_runner_open_boundary_and_dpp(tmp_path, field17):
    F = field17
    cfg = base_cfg(boundary="open", seed=9, trajectory_duration_gyroperiods=3.01)
    tr = runner.resolve_total_field_time_reference(dict(cfg, trajectory_mode="total"), F, energy_GeV=1e5)
    res = runner.run_gpu_ensemble(cfg, F, energy_GeV=1e5, time_reference=tr, output_path=str(tmp_path / "ps"),
                                  chunk_saves=4)
    s, o = oracle_for(cfg, F, res, 1e5, 1)
    pos, mom = np.asarray(res.cache["positions"]), np.asarray(res.cache["momenta"])
    assert np.array_equal(np.isnan(pos), np.isnan(o["pos"]))
    live = ~np.isnan(o["pos"])
    assert np.allclose(pos[live], o["pos"][live].astype(np.float32), rtol=1e-6)
    assert np.allclose(res.alive_fraction, o["alive_counts"] / 300.0, rtol=0, atol=0)  # integer counts: exact
    assert res.alive_fraction[-1] < 1.0
    # D_pp over the phase-space cache
    pcfg = dict(first_particle=1, n_particles_to_use=None, particle_selection="range", particle_chunk_size=100,
                lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
                lag_min_gyroperiods=None, lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None,
                n_lag_samples=6, n_energy_snapshots=5)
    t_s = res.t_s[s.save_idx - 1]
    t_norm = res.t_norm[s.save_idx - 1]
    with runner.Context(0) as ctx:
        prod = dpp.run_dpp_full(pcfg, ctx, mom, t_s, t_norm)
    acc = oracle.new_dpp_accumulators(len(prod["dpp"]["lag_steps"]))
    oracle.dpp(mom, prod["dpp"]["lag_steps"], prod["dpp"]["p0_kg_m_per_s"], acc)
    assert np.array_equal(prod["dpp"]["counts"], acc["counts"])
    assert np.allclose(prod["dpp"]["sum_delta_p2"], acc["sum_delta_p2"], rtol=1e-10)
    e = prod["energy_snapshots"]["energy_GeV"]
    assert e.shape == (5, 300) and np.all(np.isnan(e[-1]) == np.isnan(mom[-1, 0]))


def test_float32_precision_uses_production_pusher(tmp_path, field17):
    F = field17
    cfg = base_cfg(precision="Float32", seed=3)
    res = runner.run_total_field_mu_cache_ensemble(cfg, F, None, energy_GeV=10.0, chunk_saves=5)
    s, o = oracle_for(cfg, F, res, 10.0, 2)  # the same float32 dt and initial state, pushed in FP64
    assert s.T == np.float32
    mu = np.asarray(res.cache["mu"])
    assert mu.shape == o["mu"].shape
    assert np.max(np.abs(mu - o["mu"])) <= 5e-4  # FP32 fields and arithmetic over 200 pushes


def test_phase_space_dmumu_product_matches_oracle(field17):
    """run_combined_full: mu reconstructed on the device per chunk (FP64 planar
    field, rounded to the float32 mu_chunk) feeding the pair pass, vs the
    oracle's reconstruct_mu_chunk + pair accumulation on the same cache."""
    F = field17
    L = fields.box_length_m(10.0)
    g = oracle.make_geom(17, 17, 17, L)
    rng = np.random.default_rng(17)
    n = 333
    x = np.ascontiguousarray(rng.random((3, n)) * L)
    d = rng.standard_normal((3, n))
    p = np.ascontiguousarray(d / np.linalg.norm(d, axis=0) * 5.4e-18)  # ~10 GeV protons
    # test data only: dt = 2e6 s is many gyro-radians per push, so mu decorrelates
    # quickly and every bin and lag sees varied pairs
    o = oracle.run(F, g, x, p, 2.0e6, 1, 301, 301, 10, True, 1)
    pos = o["pos"].astype(np.float32)
    mom = o["mom"].astype(np.float32)
    nsave = pos.shape[0]
    t_s = 2.0e6 * 10 * np.arange(nsave)
    t_norm = t_s * 1e-7
    dcfg = dict(mu_min=0.0, mu_max=1.0, n_mu_bins=12, mu_bin_abs=True, dmumu_start_mode="sliding",
                first_particle=1, n_particles_to_use=None, particle_selection="range", particle_chunk_size=100,
                lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
                lag_min_gyroperiods=None, lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None,
                n_lag_samples=7, min_count_per_cell=5, max_pair_visits_without_allow=5e9, allow_huge=False)
    with runner.Context(0) as ctx:
        fld = runner.Field(ctx, F, L, "f64_planar")
        prod = dmumu.run_combined_full(dcfg, ctx, fld, pos, mom, t_s, t_norm)
        fld.close()
    mu_ref = oracle.reconstruct_mu_chunk(pos.astype(np.float64), mom.astype(np.float64), F[:3], g)  # (n, nsave)
    lags = prod["dmumu"]["lag_steps"]
    acc = oracle.new_dmumu_accumulators(len(lags), 12)
    for c0 in range(0, n, 100):
        oracle.dmumu(mu_ref[c0:c0 + 100].astype(np.float32), lags, prod["dmumu"]["mu_edges"], False, True, acc)
    assert acc["dmumu_counts"].sum() > 0
    assert np.array_equal(prod["dmumu"]["count_pairs_full"], acc["dmumu_counts"])
    assert np.allclose(prod["dmumu"]["sum_delta_mu2"], acc["dmumu_sum_delta2"], rtol=1e-11)
    assert np.allclose(prod["delta_mu2"]["delta_mu2_particle_mean"], acc["particle_means"], rtol=1e-11)
    assert np.array_equal(prod["delta_mu2"]["n_pairs_used"], acc["pair_counts"])
