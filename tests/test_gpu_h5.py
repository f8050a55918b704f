"""HDF5 caches and products end to end on the GPU: the runner writes its
cadence cache as an HDF5 file in the reference's layout through the
asynchronous writer, and the D_mumu / D_pp pipelines stream that file back.
Every result must equal the same run with an in-memory cache bit for bit
(same kernels, same chunk order)."""
import numpy as np
import pytest

from particle_transport_amd import cache_h5, dmumu, dpp, fields, h5, runner

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field17():
    raw = fields.generate_turbulence(17, seed=8, kmax=4, dB_over_B0=0.7)["total"]
    return [np.asfortranarray(a) for a in fields.to_si(raw)]


def base_cfg(**kw):
    # FP64 time axes (a Float32 run's axes fail the reference's own 1e-10
    # consistency check, validate_time_axes_result, time_units.jl:185), the
    # production pusher on an FP32 brick field
    cfg = dict(box_length_pc=10.0, boundary="periodic", precision="Float64", gpu_pusher="fast",
               field_storage="f32_brick", integration_steps_per_gyroperiod=50.0,
               trajectory_duration_gyroperiods=4.01, trajectory_save_interval_gyroperiods=0.2, use_cfl=False,
               cfl=0.2, eta=0.1, tOmega0_max=1.0, trajectory_time_stride=1, seed=5, n_particles=3000,
               trajectory_mode="total", injection_mode="isotropic", mu_cache_output_precision="Float32",
               trajectory_output_precision="Float32", progress_every=0, file="fields/total.h5")
    cfg.update(kw)
    return cfg


DCFG = dict(mu_min=0.0, mu_max=1.0, n_mu_bins=10, mu_bin_abs=True, dmumu_start_mode="sliding", first_particle=1,
            n_particles_to_use=2500, particle_selection="block_random", particle_seed=3, particle_block_size=100,
            particle_chunk_size=128, lag_mode="uniform_samples", lag_range_policy="fixed", lag_boundary_policy="strict",
            lag_min_gyroperiods=None, lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None,
            n_lag_samples=8, min_count_per_cell=5, max_pair_visits_without_allow=5e9, allow_huge=False,
            n_energy_snapshots=5, compute_precision="Float32")


def same(a, b):
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y, equal_nan=True), k
        elif isinstance(x, float):
            assert x == y or (np.isnan(x) and np.isnan(y)), k


def test_mu_cache_file_equals_memory_and_feeds_dmumu(tmp_path, field17):
    F = field17
    cfg = base_cfg()
    path = str(tmp_path / "cache" / "mu_cache_10_GeV.h5")
    with runner.Context(0) as ctx:
        tr = runner.resolve_total_field_time_reference(cfg, F, energy_GeV=10.0, ctx=ctx)
        a = runner.run_mu_cache_ensemble(cfg, F, None, energy_GeV=10.0, time_reference=tr, ctx=ctx, chunk_saves=7)
        b = runner.run_mu_cache_ensemble(cfg, F, path, energy_GeV=10.0, time_reference=tr, ctx=ctx, chunk_saves=7)
        mem = np.asarray(a.cache["mu"])
        assert isinstance(b.cache["mu"], h5.H5Dataset)
        assert np.array_equal(b.cache["mu"][...], mem)
        b.cache.close()
        assert cache_h5.validate_mu_cache(path).valid
        c = cache_h5.open_cache(path)
        assert c.mode == "mu" and c.mu.shape == mem.shape
        assert np.array_equal(c.alive_fraction, a.cache.meta["alive_fraction"])
        assert np.array_equal(c.t_s, a.cache.meta["t_s"])
        p_mem = dmumu.run_combined_from_mu_cache(DCFG, ctx, mem, c.t_s, c.t_norm, c.t_gyroperiods)
        p_file = dmumu.run_combined_from_mu_cache(DCFG, ctx, c.mu, c.t_s, c.t_norm, c.t_gyroperiods)
        c.file.close()
    same(p_mem["dmumu"], p_file["dmumu"])
    same(p_mem["delta_mu2"], p_file["delta_mu2"])
    assert p_file["dmumu"]["count_pairs_full"].sum() > 0
    out = str(tmp_path / "10_GeV" / "delta_mu2_dmumu_full.h5")
    cache_h5.save_combined_full_h5(out, dict(DCFG, trajectory_h5=path), p_file, backend="mu_cache")
    cache_h5.append_cache_metadata(out, "mu", path)
    got = cache_h5.read_group(out, "dmumu")
    assert np.array_equal(got["count_pairs_full"], p_file["dmumu"]["count_pairs_full"])
    with h5.H5File(out, "r") as f:
        assert f["cache_mode"] == "mu" and f["n_particles"].tolist() == [3000]
        assert f["n_particles_used"] == 2500 and f["trajectory_h5"] == path


def test_phase_space_file_feeds_dpp_and_mu_reconstruction(tmp_path, field17):
    F = field17
    cfg = base_cfg(n_particles=1500, seed=9)
    path = str(tmp_path / "phase_space_10_GeV.h5")
    with runner.Context(0) as ctx:
        tr = runner.resolve_total_field_time_reference(cfg, F, energy_GeV=10.0, ctx=ctx)
        a = runner.run_gpu_ensemble(cfg, F, energy_GeV=10.0, time_reference=tr, ctx=ctx, chunk_saves=5)
        b = runner.run_gpu_ensemble(cfg, F, energy_GeV=10.0, time_reference=tr, ctx=ctx, chunk_saves=5,
                                    output_path=path)
        pos, mom = np.asarray(a.cache["positions"]), np.asarray(a.cache["momenta"])
        assert np.array_equal(b.cache["positions"][...], pos)
        assert np.array_equal(b.cache["momenta"][...], mom)
        b.cache.close()
        assert cache_h5.validate_phase_space_cache(path).valid
        c = cache_h5.open_cache(path)
        dcfg = dict(DCFG, n_particles_to_use=1200, trajectory_h5=path)
        d_mem = dpp.run_dpp_full(dcfg, ctx, mom, c.t_s, c.t_norm, c.t_gyroperiods)
        d_file = dpp.run_dpp_full(dcfg, ctx, c.momenta, c.t_s, c.t_norm, c.t_gyroperiods)
        same(d_mem["dpp"], d_file["dpp"])
        same(d_mem["energy_snapshots"], d_file["energy_snapshots"])
        fld = runner.Field(ctx, F, fields.box_length_m(10.0), "f32_planar")
        m_mem = dmumu.run_combined_full(dcfg, ctx, fld, pos, mom, c.t_s, c.t_norm, c.t_gyroperiods)
        m_file = dmumu.run_combined_full(dcfg, ctx, fld, c.positions, c.momenta, c.t_s, c.t_norm, c.t_gyroperiods)
        fld.close()
        c.file.close()
    same(m_mem["dmumu"], m_file["dmumu"])
    assert d_file["dpp"]["counts"].sum() > 0
    out = str(tmp_path / "dpp_full.h5")
    cache_h5.save_dpp_full_h5(out, dcfg, d_file)
    with h5.H5File(out, "r") as f:
        assert f["cache_h5"] == path and f["momentum_unit"] == "kg*m/s"
        assert f["dpp/count_pairs_full"].tolist() == d_file["dpp"]["counts"].tolist()
        assert f["energy_snapshots/energy_GeV"].shape == (1200, 5)
