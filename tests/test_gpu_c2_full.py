"""The plain-GPU-path configuration at its full size, end to end through the
reference's entry points: a 256^3-cell field (float32 B and u), 1,048,576
protons at 10 GeV with isotropic injection, 200 T_g0 at 200 steps/T_g0, the
phase-space cache every 20 steps (2001 samples, 48 GB in host memory), then
D_mumu over 128 lags x 40 |mu| bins with all particles and D_pp over 128
lags with 16 energy snapshots.

At this size the oracle cannot run, so the checks are properties that hold at
any size: with a periodic box every particle stays alive and every sample is
finite, so each lag has exactly N (nsave - lag) pairs, all of them binned
(|mu| <= 1, last edge inclusive); the per-particle means are over N
particles; the energy snapshots stay at the injection energy to the level the
motional electric field allows; and the results do not depend on the particle
chunk size (the reference merges chunks with Chan's formula)."""
import math

import numpy as np
import pytest

from particle_transport_amd import dmumu, dpp, fields, runner

pytestmark = pytest.mark.gpu

N = 1_048_576


@pytest.fixture(scope="module")
def field257():
    raw = fields.generate_turbulence(257, seed=2, kmin=1.0, kmax=16.0, dB_over_B0=0.5, dtype="float32")["total"]
    return [np.asfortranarray(a, dtype=np.float32) for a in fields.to_si(raw)]


def test_c2_full_size_end_to_end(field257):
    F = field257
    # FP64 time axes (a Float32 run's axes fail the reference's own uniformity
    # check, so no product could be computed from its cache); the production
    # pusher on an FP32 brick field
    cfg = dict(box_length_pc=10.0, boundary="periodic", precision="Float64", gpu_pusher="fast",
               field_storage="f32_brick", integration_steps_per_gyroperiod=200.0,
               trajectory_duration_gyroperiods=200.0, trajectory_save_interval_gyroperiods=0.1, use_cfl=False,
               cfl=0.2, eta=0.1, tOmega0_max=1.0, trajectory_time_stride=1, seed=42, n_particles=N,
               trajectory_mode="total", injection_mode="isotropic", injection_rng="device",
               trajectory_output_precision="Float32", progress_every=0)
    with runner.Context(0) as ctx:
        tr = runner.resolve_total_field_time_reference(cfg, F, energy_GeV=10.0, ctx=ctx)
        res = runner.run_gpu_ensemble(cfg, F, energy_GeV=10.0, time_reference=tr, ctx=ctx)
        s = res.save_time
        assert s.trajectory_time_stride == 20
        nsave = res.nsave
        assert nsave == (res.trajectory_time.n_integration_steps - 1) // 20 + 1
        assert 1990 <= nsave <= 2010
        assert np.all(res.alive_fraction == 1.0)
        pos, mom = res.cache["positions"], res.cache["momenta"]
        assert pos.shape == mom.shape == (nsave, 3, N)
        assert np.isfinite(mom[-1]).all() and np.isfinite(pos[nsave // 2]).all()
        meta = res.cache.meta
        fld = runner.Field(ctx, F, fields.box_length_m(10.0), "f32_planar")
        dcfg = dict(mu_min=0.0, mu_max=1.0, n_mu_bins=40, mu_bin_abs=True, dmumu_start_mode="sliding",
                    first_particle=1, n_particles_to_use=None, particle_selection="range", particle_chunk_size=128,
                    lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
                    lag_min_gyroperiods=None, lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None,
                    n_lag_samples=128, min_count_per_cell=20, allow_huge=True, n_energy_snapshots=16)
        prod = dmumu.run_combined_full(dcfg, ctx, fld, pos, mom, meta["t_s"], meta["t_norm"], meta["t_gyroperiods"])
        fld.close()
        dp = dpp.run_dpp_full(dcfg, ctx, mom, meta["t_s"], meta["t_norm"], meta["t_gyroperiods"])
        # the same D_pp with chunks of 2048 particles: only the merge tree's rounding differs
        dp2 = dpp.run_dpp_full(dict(dcfg, particle_chunk_size=2048), ctx, mom, meta["t_s"], meta["t_norm"],
                               meta["t_gyroperiods"])
        res.cache.close()
    lags = np.asarray(prod["dmumu"]["lag_steps"])
    assert 100 <= len(lags) <= 128
    pairs = N * (nsave - lags)
    counts = np.asarray(prod["dmumu"]["count_pairs_full"])
    assert counts.shape == (len(lags), 40)
    assert np.array_equal(counts.sum(axis=1), pairs)  # every pair binned
    assert np.array_equal(prod["delta_mu2"]["n_pairs_used"], pairs)
    assert np.array_equal(prod["delta_mu2"]["n_particles_used"], np.full(len(lags), N))
    d2 = np.asarray(prod["delta_mu2"]["delta_mu2_particle_mean"])
    assert np.all(np.isfinite(d2)) and np.all((d2 > 0) & (d2 <= 4.0))
    assert np.all(np.isfinite(prod["dmumu"]["D_mumu_centered_norm"]))  # every cell has >= 20 pairs
    dlags = np.asarray(dp["dpp"]["lag_steps"])
    assert np.array_equal(dp["dpp"]["counts"], N * (nsave - dlags))
    assert np.all(np.isfinite(dp["dpp"]["dpp_centered_norm"]))
    e = np.asarray(dp["energy_snapshots"]["energy_GeV"])
    assert e.shape == (16, N) and np.isfinite(e).all()
    assert np.max(np.abs(e / 10.0 - 1.0)) < 1e-2
    assert np.array_equal(dp2["dpp"]["counts"], dp["dpp"]["counts"])
    for k in ("sum_delta_p2", "sum_delta_p_norm2"):
        assert np.allclose(dp2["dpp"][k], dp["dpp"][k], rtol=1e-11, atol=0), k
    assert math.isfinite(float(dp["dpp"]["p0_kg_m_per_s"]))
