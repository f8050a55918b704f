"""HDF5 files in the reference's layouts (libpt_h5.so, particle_transport_amd.h5
and cache_h5), on the CPU.

The layout is checked with HDF5's own h5dump / h5ls, independently of this
build's reader: a numpy (nsave, 3, n) array must appear as the dataspace
(nsave, 3, n), which HDF5.jl reads as the Julia (n, 3, nsave) array the
reference writes (create_phase_space_writer, phase_space_gpu_runner.jl:406-411;
the reference's Python reader expects momenta shaped (nsave, 3, n),
plot_energy_distribution.py:109-112).  Scalars must keep the reference's two
forms: Float64[x] (a length-1 vector) and Int(x) (a true scalar)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from particle_transport_amd import cache_h5, dmumu, dpp, h5, runner, time_units

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
H5LS = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
needs_tools = pytest.mark.skipif(not (os.path.exists(H5DUMP) and os.path.exists(H5LS)),
                                 reason="HDF5 command-line tools are not installed")


def dump(*args):
    r = subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True)
    return r.stdout


def ls(path):
    return subprocess.run([H5LS, "-r", "-v", path], capture_output=True, text=True, check=True).stdout


@needs_tools
def test_cadence_dataset_layout_and_slabs(tmp_path):
    path = str(tmp_path / "phase.h5")
    nsave, n = 7, 5
    a = (np.arange(nsave * 3 * n, dtype=np.float32) * 0.5).reshape(nsave, 3, n)
    with h5.H5File(path, "w") as f:
        f.create_dataset("positions", (nsave, 3, n), np.float32)
        f.write_slab("positions", 0, a[:3])  # two chunks of save indices, as the writer thread does
        f.write_slab("positions", 3, a[3:])
    out = dump("-d", "positions", "-s", "4,2,1", "-c", "1,1,3", path)
    assert "DATASPACE  SIMPLE { ( 7, 3, 5 ) / ( 7, 3, 5 ) }" in out
    assert "H5T_IEEE_F32LE" in out
    vals = [float(v) for v in out.split("(4,2,1):")[1].split("\n")[0].split(",")]
    assert vals == [float(x) for x in a[4, 2, 1:4]]
    with h5.H5File(path, "r") as f:
        ds = f.dataset("positions")
        assert ds.shape == (nsave, 3, n) and ds.dtype == np.float32
        assert np.array_equal(ds[...], a)
        assert np.array_equal(ds[2:5], a[2:5])
        assert np.array_equal(ds.read_columns([4, 0, 1, 3]), a[..., [4, 0, 1, 3]])
        assert np.array_equal(ds.read_box((1, 0, 2), (3, 2, 2), np.float64), a[1:4, 0:2, 2:4].astype(np.float64))


@needs_tools
def test_scalar_string_and_bool_forms(tmp_path):
    path = str(tmp_path / "meta.h5")
    with h5.H5File(path, "w") as f:
        f["energy_GeV"] = np.array([10.0])  # Float64[energy_GeV]
        f["n_particles"] = np.array([4096], dtype=np.int64)  # Int[n]
        f["particle_chunk_size"] = 128  # Int(x): a true scalar
        f["box_length_pc"] = 200.0
        f["trajectory_mode"] = "total"
        f["empty"] = ""
        f["exact_final_state_stored"] = False
        f["mu_bin_abs"] = True
        f["modes"] = ["alfven", "fast"]
        f["dmumu/count_pairs_full"] = np.arange(6, dtype=np.int64).reshape(3, 2)  # Julia (2, 3)
    text = ls(path)
    assert "/energy_GeV              Dataset {1/1}" in text
    assert "/particle_chunk_size     Dataset {SCALAR}" in text
    assert "5-byte null-terminated UTF-8 string" in text
    assert "variable-length null-terminated UTF-8 string" in text
    assert 'FALSE            = 0' in text and 'TRUE             = 1' in text
    assert "/dmumu/count_pairs_full  Dataset {3/3, 2/2}" in text
    with h5.H5File(path, "r") as f:
        assert isinstance(f["energy_GeV"], np.ndarray) and f["energy_GeV"].tolist() == [10.0]
        assert f["n_particles"].dtype == np.int64 and f["n_particles"].tolist() == [4096]
        assert f["particle_chunk_size"] == 128 and isinstance(f["particle_chunk_size"], int)
        assert f["box_length_pc"] == 200.0
        assert f["trajectory_mode"] == "total" and f["empty"] == ""
        assert f["exact_final_state_stored"] is False and f["mu_bin_abs"] is True
        assert f["modes"] == ["alfven", "fast"]
        assert np.array_equal(f["dmumu/count_pairs_full"], np.arange(6).reshape(3, 2))
        assert "dmumu" in f and "dmumu/count_pairs_full" in f and "dmumu/nope" not in f and "x/y" not in f
        # h5_scalar_value_or / h5_scalar_bool_or accept both forms
        assert h5.scalar_value_or(f, "energy_GeV", None) == 10.0
        assert h5.scalar_value_or(f, "particle_chunk_size", None) == 128
        assert h5.scalar_value_or(f, "missing", "d") == "d"
        assert h5.scalar_bool_or(f, "mu_bin_abs", False) is True
        assert h5.scalar_bool_or(f, "n_particles", False) is True
        assert h5.scalar_bool_or(f, "trajectory_mode", True) is True  # not a flag: default


def test_errors_are_reported(tmp_path):
    path = str(tmp_path / "e.h5")
    with h5.H5File(path, "w") as f:
        f.create_dataset("mu", (4, 3), np.float32)
        with pytest.raises(h5.H5Error, match="already exists"):
            f.create_dataset("mu", (4, 3), np.float32)
        with pytest.raises(h5.H5Error, match="box outside dataset 'mu'"):
            f.write_slab("mu", 3, np.zeros((2, 3), np.float32))
        with pytest.raises(h5.H5Error, match="no dataset 'nope'"):
            f.write_slab("nope", 0, np.zeros((1, 3), np.float32))
        f["s"] = "x"
        with pytest.raises(h5.H5Error, match="not a numeric dataset"):
            f.dataset("s")
    with pytest.raises(h5.H5Error, match="cannot open"):
        h5.H5File(str(tmp_path / "missing.h5"), "r")


@pytest.mark.parametrize("selection", ["range", "random", "block_random", "reversed"])
def test_particle_groups_follow_the_selection(tmp_path, selection):
    nsave, n, chunk = 9, 1000, 64
    rng = np.random.default_rng(5)
    a = rng.standard_normal((nsave, 3, n)).astype(np.float32)
    path = str(tmp_path / "g.h5")
    with h5.H5File(path, "w") as f:
        f.create_dataset("momenta", a.shape, np.float32)
        f.write_slab("momenta", 0, a)
    cfg = dict(first_particle=3, n_particles_to_use=500, particle_selection=selection.replace("reversed", "range"),
               particle_seed=11, particle_block_size=50)
    pidx0 = dmumu.build_particle_indices(n, cfg) - 1
    if selection == "reversed":
        pidx0 = pidx0[::-1].copy()
    with h5.H5File(path, "r") as f:
        for src in (f.dataset("momenta"), h5.ArrayDataset(a)):
            for budget in (1 << 30, 20_000, 1):  # one span, several spans, per-chunk runs
                for reuse in (False, True):
                    seen = []
                    for (block,), chunks in h5.iter_particle_groups([src], pidx0, chunk, budget, reuse=reuse):
                        for c0, local in chunks:
                            want = pidx0[c0:c0 + chunk]
                            assert np.array_equal(block[..., local], a[..., want])
                            seen.append(c0)
                    assert seen == list(range(0, len(pidx0), chunk))


def _setup(cfg, F):
    tr = time_units.campaign_time_reference(5.0e-10, 0.3, source_path="synthetic")
    return runner._trajectory_setup(cfg, F, 10.0, tr, "test")


CFG = dict(box_length_pc=10.0, boundary="periodic", precision="Float64", integration_steps_per_gyroperiod=50.0,
           trajectory_duration_gyroperiods=2.01, trajectory_save_interval_gyroperiods=0.2, use_cfl=False,
           cfl=0.2, eta=0.1, tOmega0_max=1.0, trajectory_time_stride=1, seed=17, n_particles=64,
           trajectory_mode="total", injection_mode="fixed-mu", injection_mu0=0.6, file="fields.h5")


def _write_cache(path, kind, s, nsave, n, alive=None):
    alive = np.ones(s.nsteps) if alive is None else alive
    with h5.H5File(path, "w") as f:
        if kind == "mu":
            f.create_dataset("mu", (nsave, n), np.float32)
            f.write_slab("mu", 0, np.random.default_rng(1).uniform(-1, 1, (nsave, n)).astype(np.float32))
        else:
            for name in ("positions", "momenta"):
                f.create_dataset(name, (nsave, 3, n), np.float32)
        cache_h5.finalize_cache(f, CFG, s, 10.0, alive, kind, np.float32)


def test_group_plans_cover_the_selection_in_order():
    """plan_particle_groups / plan_chunks: whole chunks in selection order,
    spans while the column span fits the cap and indices ascend, runs
    otherwise; every selected index appears once, at its chunk's offset."""
    assert h5.plan_particle_groups(np.zeros(0, np.int64), 4, 100) == []
    pidx0 = np.arange(10, 30)
    assert h5.plan_particle_groups(pidx0, 4, 100) == [("span", 0, 20, 10, 29)]
    plans = h5.plan_particle_groups(pidx0, 4, 9)  # two chunks (8 columns) per span
    assert [p[:3] for p in plans] == [("span", 0, 8), ("span", 8, 16), ("span", 16, 20)]
    desc = pidx0[::-1].copy()
    assert all(p[0] == "runs" for p in h5.plan_particle_groups(desc, 4, 100))
    rng = np.random.default_rng(3)
    sel = np.sort(rng.choice(1000, 300, replace=False))
    for cap in (1, 50, 10_000):
        plans = h5.plan_particle_groups(sel, 16, cap)
        seen = []
        for plan in plans:
            kind, g0, g1, lo, hi = plan
            if kind == "span":
                assert hi - lo + 1 <= max(cap, sel[min(g0 + 16, g1) - 1] - sel[g0] + 1)
            for c0, local in h5.plan_chunks(plan, sel, 16):
                want = sel[c0:min(c0 + 16, g1)]
                got = local + lo if kind == "span" else sel[g0:g1][local]
                assert np.array_equal(got, want)
                seen.append(c0)
        assert seen == list(range(0, len(sel), 16))


def test_read_box_into_given_buffers():
    """A box large enough to be copied on several threads, read into a fresh
    array and into a caller's buffer, equals the slice; a buffer of the wrong
    shape is refused."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((20, 3, 70_000)).astype(np.float32)
    src = h5.ArrayDataset(a)
    start, count = (0, 0, 1_234), (20, 3, 60_001)
    want = a[:, :, 1_234:61_235]
    got = src.read_box(start, count)
    assert got.flags.c_contiguous and np.array_equal(got, want)
    out = np.full(count, np.nan, np.float32)
    assert src.read_box(start, count, out=out) is out and np.array_equal(out, want)
    wide = src.read_box(start, count, dtype=np.float64)
    assert wide.dtype == np.float64 and np.array_equal(wide, want.astype(np.float64))
    with pytest.raises(h5.H5Error):
        src.read_box(start, count, out=np.empty((20, 3, 60_000), np.float32))


def test_cache_metadata_and_validators(tmp_path):
    F = [np.zeros((9, 9, 9), order="F") for _ in range(6)]
    s = _setup(CFG, F)
    nsave = len(s.save_idx)
    mu_path, ps_path = str(tmp_path / "mu.h5"), str(tmp_path / "ps.h5")
    _write_cache(mu_path, "mu", s, nsave, 64)
    _write_cache(ps_path, "phase", s, nsave, 64)
    assert cache_h5.validate_mu_cache(mu_path).valid
    assert cache_h5.validate_phase_space_cache(ps_path).valid
    assert cache_h5.validate_cache(mu_path, ":mu").valid
    with h5.H5File(mu_path, "r") as f:
        assert f["cache_mode"] == "mu" and f["cache_output_precision"] == "Float32" and f["mu_unit"] == "dimensionless"
        assert f["energy_GeV"].tolist() == [10.0] and f["n_particles"].tolist() == [64]
        assert f["analysis_sample_count"].tolist() == [nsave]
        assert f["trajectory_time_stride"].tolist() == [s.st.trajectory_time_stride]
        assert f["exact_final_state_stored"] is False
        assert f["time_reference_name"] == time_units.REFERENCE_GYROPERIOD_NAME
        assert f["time_reference_source_path"] == "synthetic"
        assert f["trajectory_field_source_path"] == "fields.h5"
        assert f["injection_mode"] == "fixed_mu" and f["injection_mu0"].tolist() == [0.6]
        assert f["t_gyroperiods"].shape == (nsave,)
        assert math.isclose(f["analysis_cache_duration_gyroperiods"][0], f["t_gyroperiods"][-1] - f["t_gyroperiods"][0])
    with h5.H5File(ps_path, "r") as f:
        assert f["trajectory_output_precision"] == "Float32" and f["position_unit"] == "m"
        assert f["momentum_unit"] == "kg*m/s" and "cache_mode" not in f
    # the reference's failure reasons
    r = cache_h5.validate_mu_cache(str(tmp_path / "missing.h5"))
    assert not r.valid and r.reason == "legacy_metadata_missing"
    bad = str(tmp_path / "bad.h5")
    with h5.H5File(bad, "w") as f:
        f.create_dataset("mu", (nsave + 1, 64), np.float32)
        cache_h5.finalize_cache(f, CFG, s, 10.0, np.ones(s.nsteps), "mu", np.float32)
    r = cache_h5.validate_mu_cache(bad)
    assert not r.valid and r.reason == "time_length_mismatch" and "t_s length mismatch" in r.message
    assert not cache_h5.validate_phase_space_cache(mu_path).valid  # no positions dataset
    c = cache_h5.open_cache(mu_path)
    assert c.mode == "mu" and c.mu.shape == (nsave, 64) and len(c.t_s) == nsave
    c.file.close()


def _synthetic_dmumu_products(cfg, nsave, n):
    rng = np.random.default_rng(3)
    t_s = np.arange(nsave) * 0.5
    t_norm = t_s * 0.3
    t_gp = t_norm / (2 * math.pi)
    u = dmumu._setup_pass(cfg, t_s, t_norm, t_gp, nsave, n, "Mu-cache time axes")
    nl, nb = len(u.lag_steps), len(u.centers)
    acc = u.acc
    acc.dmumu_counts[:] = rng.integers(0, 50, (nl, nb))
    acc.dmumu_sum_delta[:] = rng.standard_normal((nl, nb))
    acc.dmumu_sum_delta2[:] = rng.uniform(0, 3, (nl, nb))
    acc.particle_counts[:] = rng.integers(2, 9, nl)
    acc.particle_means[:] = rng.uniform(0, 1, nl)
    acc.particle_m2s[:] = rng.uniform(0, 1, nl)
    acc.pair_sum_squares[:] = rng.uniform(0, 5, nl)
    acc.pair_counts[:] = rng.integers(1, 99, nl)
    return dmumu._finish_products(cfg, acc, u.lag_steps, u.grid, u.tau_s, u.tau_norm, u.t_s, u.t_norm, u.edges,
                                  u.centers, u.visits, u.pidx)


DCFG = dict(first_particle=1, n_particles_to_use=None, particle_selection="range", particle_seed=1,
            particle_block_size=8, lag_mode="uniform_samples", n_lag_samples=6, lag_step_stride=1,
            max_lag_steps=None, min_count_per_cell=5, mu_min=0.0, mu_max=1.0, n_mu_bins=4, mu_bin_abs=True,
            particle_chunk_size=16, trajectory_h5="cache/mu_cache_10_GeV.h5", compute_precision="Float32",
            box_length_pc=10.0, allow_huge=True)


@needs_tools
def test_dmumu_product_file(tmp_path):
    prod = _synthetic_dmumu_products(DCFG, 40, 50)
    path = str(tmp_path / "10_GeV" / "delta_mu2_dmumu_full.h5")
    cache_h5.save_combined_full_h5(path, DCFG, prod, backend="mu_cache")
    nl, nb = prod["dmumu"]["count_pairs_full"].shape
    text = ls(path)
    assert f"/dmumu/count_pairs_full  Dataset {{{nl}/{nl}, {nb}/{nb}}}" in text  # Julia (n_bins, n_lags)
    assert "/dmumu/D_mumu_centered_tau_average_norm_count_weighted" in text
    assert "/delta_mu2/n_pairs_used" in text
    got = cache_h5.read_group(path, "dmumu")
    for k in ("count_pairs_full", "sum_delta_mu", "sum_delta_mu2", "D_mumu_centered_norm", "mu_edges"):
        assert np.array_equal(got[k], np.asarray(prod["dmumu"][k]), equal_nan=True), k
    assert np.array_equal(got["D_mumu_raw_tau_average_norm"], prod["dmumu"]["tau_avg_raw_norm"], equal_nan=True)
    df = cache_h5.read_group(path, "delta_mu2")
    assert np.array_equal(df["delta_mu2_pair_mean"], prod["delta_mu2"]["delta_mu2_pair_mean"], equal_nan=True)
    with h5.H5File(path, "r") as f:
        assert f["compute_backend"] == "mu_cache" and f["dmumu_start_mode"] == "sliding"
        assert f["mu_bin_abs"] is True and f["mu_bin_coordinate"] == "abs(mu(t))"
        assert f["n_particles_used"] == 50 and f["first_particle"] == 1 and f["last_particle"] == 50
        assert np.array_equal(f["particle_indices"], np.arange(1, 51))
        assert f["estimated_pair_visits"] == str(prod["estimated_pair_visits"])
        assert f["max_lag_steps"] == -1 and f["source_cache_uniform_time_axis"] is True
        assert f["unique_lag_count"].tolist() == [nl]
        assert f["lag_grid_source"] == str(prod["lag_grid"].lag_source).lstrip(":")
        assert f["source_cache_identity"] == "cache/mu_cache_10_GeV.h5"


@needs_tools
def test_dpp_product_file_and_cache_metadata_copy(tmp_path):
    F = [np.zeros((9, 9, 9), order="F") for _ in range(6)]
    s = _setup(CFG, F)
    nsave = len(s.save_idx)
    cache = str(tmp_path / "phase_space_10_GeV.h5")
    _write_cache(cache, "phase", s, nsave, 64)
    rng = np.random.default_rng(9)
    lag_grid = time_units.resolve_lag_grid(DCFG, np.asarray(s.t_gp[s.save_idx - 1], dtype=np.float64))
    nl = len(lag_grid.lag_steps)
    acc = dpp.new_accumulators(nl)
    acc.counts[:] = rng.integers(1, 100, nl)
    acc.sum_delta_p[:] = rng.standard_normal(nl)
    acc.sum_delta_p2[:] = rng.uniform(0, 2, nl)
    acc.sum_delta_p_norm[:] = rng.standard_normal(nl)
    acc.sum_delta_p_norm2[:] = rng.uniform(0, 2, nl)
    tau_s = rng.uniform(1, 2, nl)
    arrays = dpp.compute_dpp_arrays(acc, tau_s, lag_grid.tau_norm)
    energy = rng.uniform(1, 2, (5, 64))
    prod = dict(dpp=dict(p0_kg_m_per_s=1.5, counts=acc.counts, sum_delta_p=acc.sum_delta_p,
                         sum_delta_p2=acc.sum_delta_p2, sum_delta_p_norm=acc.sum_delta_p_norm,
                         sum_delta_p_norm2=acc.sum_delta_p_norm2, lag_steps=lag_grid.lag_steps, tau_s=tau_s,
                         tau_norm=lag_grid.tau_norm, tau_gyroperiods=lag_grid.tau_gyroperiods, **arrays),
                energy_snapshots=dict(snapshot_step_index=np.array([1, 3, 5, 8, 11]), snapshot_t_s=np.arange(5.0),
                                      snapshot_t_norm=np.arange(5.0), snapshot_t_gyroperiods=np.arange(5.0),
                                      energy_GeV=energy, particle_indices=np.arange(1, 65)),
                lag_grid=lag_grid)
    path = str(tmp_path / "dpp_full.h5")
    cache_h5.save_dpp_full_h5(path, dict(DCFG, cache_h5=cache), prod)
    text = ls(path)
    assert "/energy_snapshots/energy_GeV Dataset {64/64, 5/5}" in text  # Julia (k, N_sel)
    got = cache_h5.read_group(path, "dpp")
    assert np.array_equal(got["D_pp_centered_norm"], arrays["dpp_centered_norm"], equal_nan=True)
    assert np.array_equal(got["sum_delta_p_over_p0_2"], acc.sum_delta_p_norm2)
    assert got["p0_kg_m_per_s"].tolist() == [1.5]
    es = cache_h5.read_group(path, "energy_snapshots")
    assert np.array_equal(es["energy_GeV"].T, energy)
    with h5.H5File(path, "r") as f:
        assert f["cache_mode"] == "phase_space" and f["cache_h5"] == cache
        # copy_cache_metadata! (compute_dpp_full.jl:332-340) brought the cache's metadata along
        assert f["energy_GeV"].tolist() == [10.0] and f["momentum_unit"] == "kg*m/s"
        assert f["time_reference_source_path"] == "synthetic"
        assert "position_unit" not in f  # not in the reference's copy list
    # append_cache_metadata! on a D_mumu product from a mu cache
    mu_cache = str(tmp_path / "mu_cache_10_GeV.h5")
    _write_cache(mu_cache, "mu", s, nsave, 64)
    prod2 = _synthetic_dmumu_products(DCFG, 40, 50)
    out2 = str(tmp_path / "dm.h5")
    cache_h5.save_combined_full_h5(out2, DCFG, prod2, backend="mu_cache")
    cache_h5.append_cache_metadata(out2, ":mu", mu_cache)
    with h5.H5File(out2, "r") as f:
        assert f["cache_mode"] == "mu" and f["cache_dataset"] == "mu" and f["cache_h5"] == mu_cache
        assert f["mu_unit"] == "dimensionless" and f["n_particles"].tolist() == [64]
        assert f["box_length_pc"] == 10.0  # already present: not overwritten by the copy


def test_float32_time_axes_fail_the_reference_consistency_check(tmp_path):
    """A Float32 run computes t_gyroperiods = t_norm / T(2 pi) in Float32
    (phase_space_gpu_runner.jl:500-502); stored as Float64 they differ from
    t_norm / 2 pi by ~1e-8 relative, more than validate_time_axes_result's
    rtol = 1e-10 (time_units.jl:185), so the reference's validator rejects its
    own Float32 caches.  The port keeps that behaviour."""
    F = [np.zeros((9, 9, 9), order="F") for _ in range(6)]
    s = _setup(dict(CFG, precision="Float32"), F)
    path = str(tmp_path / "mu32.h5")
    _write_cache(path, "mu", s, len(s.save_idx), 8)
    r = cache_h5.validate_mu_cache(path)
    assert not r.valid and r.reason == "inconsistent_time_units"
    assert "t_norm and t_gyroperiods are inconsistent" in r.message


@needs_tools
def test_field_file_loader_reads_julia_arrays(tmp_path):
    """load_static_fields from an HDF5 field file (phase_space_gpu_runner.jl:95-123):
    each dataset is a Julia (nx, ny, nz) array, which HDF5 stores as the
    dataspace (nz, ny, nx).  h5dump fixes where one node sits in the file; the
    loader must return it at [i, j, k] in SI units, and field_subset must keep
    the first nodes of each axis."""
    from particle_transport_amd import fields
    nx, ny, nz = 5, 4, 3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    jl = [np.asfortranarray(1000.0 * c + 100 * i + 10 * j + k) for c in range(6)]
    path = str(tmp_path / "field.h5")
    with h5.H5File(path, "w") as f:
        for name, a in zip(fields.B_PATHS + fields.V_PATHS, jl):
            f[name] = np.ascontiguousarray(a.T)  # numpy (nz, ny, nx) = Julia (nx, ny, nz)
    out = dump("-d", "j_velocity", "-s", "2,1,3", "-c", "1,1,1", path)
    assert "DATASPACE  SIMPLE { ( 3, 4, 5 ) / ( 3, 4, 5 ) }" in out
    assert float(out.split("(2,1,3):")[1].split("\n")[0]) == 4000.0 + 100 * 3 + 10 * 1 + 2
    cfg = dict(file=path, velocity_unit_in_m_per_s=1e3)
    got = fields.load_static_fields(cfg)
    want = fields.load_static_fields(cfg, arrays=jl)
    for g, w, c in zip(got, want, range(6)):
        assert g.shape == (nx, ny, nz) and g.flags.f_contiguous
        assert np.array_equal(g, w)
        assert g[3, 1, 2] == (1000.0 * c + 312) * (1e-10 if c < 3 else 1e3)
    sub = fields.load_static_fields(dict(cfg, field_subset=(3, 2, 2)))
    assert all(np.array_equal(s, w[:3, :2, :2]) for s, w in zip(sub, want))
    # a subset larger than the dataset is an error, as data[1:nx,1:ny,1:nz] is
    # in the reference (phase_space_gpu_runner.jl:100-104), from a file or arrays
    from particle_transport_amd.time_units import PipelineError
    with pytest.raises(PipelineError, match="exceeds"):
        fields.load_static_fields(dict(cfg, field_subset=(3, 5, 2)))
    with pytest.raises(PipelineError, match="exceeds"):
        fields.load_static_fields(dict(cfg, field_subset=(6, 2, 2)), arrays=jl)
