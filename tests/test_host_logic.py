"""Host-side logic: Julia range/round semantics, bins, particle selection,
finalisation closed forms, the synthetic field generator and injection."""
import math
from fractions import Fraction
from types import SimpleNamespace

import numpy as np
import pytest

from particle_transport_amd import dmumu, dpp, fields, injection
from particle_transport_amd.julia_compat import colon, range_int_length, range_length, rat, round_int
from particle_transport_amd.time_units import PipelineError, resolve_lag_grid


# ------------------------------------------------------------ Julia ranges

def test_rat_matches_continued_fraction_results():
    assert rat(0.25) == (1, 4)
    assert rat(0.1) == (1, 10)
    assert rat(0.3) == (3, 10)
    assert rat(1.0 / 3.0) == (1, 3)
    assert rat(3.0e17)[1] == 0  # beyond maxintfloat(Float32): no rational form


def test_range_length_elements_are_exact_rationals_rounded_once():
    e = range_length(0.0, 1.0, 41)
    assert all(e[i] == i / 40 for i in range(41))  # i/40 is the once-rounded rational
    e = range_length(-1.0, 1.0, 25)
    assert all(e[i] == float(Fraction(2 * i - 24, 24)) for i in range(25))
    assert range_length(0.3, 0.7, 3)[1] == 0.5


def test_colon_lengths_match_julia():
    assert len(colon(0.0, 0.25, 5.0)) == 21
    assert len(colon(0.0, 0.30, 4.8)) == 17
    assert colon(0.0, 0.30, 4.8)[-1] == 4.8
    assert colon(0.0, 0.1, 0.3) == [0.0, 0.1, 0.2, 0.3]  # rational path hits 0.3 exactly
    assert len(colon(0.25, 0.25, 1.0 + 0.25e-12)) == 4


def test_round_int_is_half_even():
    assert [round_int(v) for v in (0.5, 1.5, 2.5, -0.5, 2.4999999)] == [0, 2, 2, 0, 2]
    with pytest.raises(ValueError):
        round_int(float("nan"))


def test_uniform_sample_lag_steps_dedupe_in_order():
    # unique(round.(Int, range(1, 10, length=7))) in Julia: 1.0 2.5 4.0 5.5 7.0 8.5 10.0 -> 1 2 4 6 7 8 10
    cfg = dict(lag_mode="uniform_samples", min_lag_steps=1, max_lag_steps=10, n_lag_samples=7)
    assert dmumu.build_selected_lag_steps(100, cfg) == [1, 2, 4, 6, 7, 8, 10]
    assert range_int_length(1, 10, 7)[1] == 2.5


def test_log_lag_grid_maps_through_nearest_and_dedupe():
    t = np.arange(501) * 0.1
    cfg = dict(lag_mode="log_samples", lag_range_policy="fixed", lag_boundary_policy="strict", lag_min_gyroperiods=0.1,
               lag_max_gyroperiods=50.0, n_lag_samples=64)
    g = resolve_lag_grid(cfg, t)
    assert g.lag_steps[0] == 1 and g.lag_steps[-1] == 500
    assert np.all(np.diff(g.lag_steps) > 0)
    assert g.unique_lag_count + g.duplicate_lag_mapping_count == 64
    assert g.duplicate_lag_mapping_count > 0  # small lags collapse onto the cadence


# ------------------------------------------------------------------- bins

def test_mu_edges_and_bin_rules():
    e, c = dmumu.build_mu_edges_full(dict(mu_min=0.0, mu_max=1.0, n_mu_bins=20, mu_bin_abs=True))
    assert len(e) == 21 and e[0] == 0.0 and e[-1] == 1.0 and e[7] == 7 / 20
    assert np.allclose(c, (np.arange(20) + 0.5) / 20)
    assert dmumu.mu_bin_index_full(1.0, e) == 20
    assert dmumu.mu_bin_index_full(0.35, e) == 8
    assert dmumu.mu_bin_index_full(-0.01, e) == 0
    with pytest.raises(PipelineError):
        dmumu.build_mu_edges_full(dict(mu_min=-1.0, mu_max=1.0, n_mu_bins=20, mu_bin_abs=True))


def _ceil_f(e):
    f = np.float32(e)
    if float(f) < e:
        f = np.nextafter(f, np.float32(np.inf))
    return f


def test_float_thresholds_equal_fp64_edge_comparisons():
    """The D_mumu kernel compares float32 mu against float thresholds; for
    every float v, v >= ceil_f(e) <=> (double)v >= e, and v <= floor_f(e) <=>
    (double)v <= e.  Checked exhaustively near each edge."""
    edges = np.asarray(range_length(0.0, 1.0, 41))
    for e in edges:
        t = _ceil_f(e)
        v = np.float32(e)
        cands = [v]
        for _ in range(3):
            cands.append(np.nextafter(cands[-1], np.float32(np.inf)))
        cands2 = [v]
        for _ in range(3):
            cands2.append(np.nextafter(cands2[-1], np.float32(-np.inf)))
        for c in cands + cands2:
            assert (c >= t) == (float(c) >= e)


# -------------------------------------------------------------- selection

def test_particle_indices_range_and_block_random_structure():
    cfg = dict(first_particle=5, n_particles_to_use=10, particle_selection="range")
    assert list(dmumu.build_particle_indices(100, cfg)) == list(range(5, 15))
    cfg = dict(first_particle=1, n_particles_to_use=None, particle_selection="range")
    assert len(dmumu.build_particle_indices(7, cfg)) == 7
    cfg = dict(first_particle=1, n_particles_to_use=300, particle_selection="block_random", particle_seed=1,
               particle_block_size=128)
    idx = dmumu.build_particle_indices(1000, cfg)
    assert len(idx) <= 300 and np.all(np.diff(idx) > 0) and idx.min() >= 1 and idx.max() <= 1000
    cfg = dict(first_particle=1, n_particles_to_use=50, particle_selection="random", particle_seed=3)
    idx = dmumu.build_particle_indices(200, cfg)
    assert len(idx) == 50 and len(set(idx)) == 50 and np.all(np.diff(idx) > 0)
    with pytest.raises(PipelineError):
        dmumu.build_particle_indices(10, dict(first_particle=11, n_particles_to_use=1, particle_selection="range"))


def test_pair_visit_estimate():
    assert dmumu.estimate_pair_visits(100, 10, [1, 10]) == 10 * (99 + 90)
    assert dmumu.estimate_pair_visits(100, 10, [1, 10], "injection") == 20


# ----------------------------------------------------------- finalisation

def test_dmumu_arrays_closed_form_and_min_count():
    cfg = dict(min_count_per_cell=3)
    counts = np.array([[4, 2]])
    s1 = np.array([[0.4, 1.0]])
    s2 = np.array([[0.2, 1.0]])
    out = dmumu.compute_dmumu_arrays_full(counts, s1, s2, [2.0], [0.5], cfg)
    mean_d, mean_d2, drift, raw_n, cen_n, raw_s, cen_s = out
    assert mean_d[0, 0] == pytest.approx(0.1) and mean_d2[0, 0] == pytest.approx(0.05)
    assert drift[0, 0] == pytest.approx(0.1 / 0.5)
    assert raw_n[0, 0] == pytest.approx(0.05 / 1.0) and cen_n[0, 0] == pytest.approx((0.05 - 0.01) / 1.0)
    assert raw_s[0, 0] == pytest.approx(0.05 / 4.0)
    assert all(np.isnan(a[0, 1]) for a in out)
    u, w = dmumu.average_over_tau_full(np.array([[1.0, np.nan], [3.0, 2.0]]), np.array([[5, 9], [15, 9]]), cfg)
    assert u[0] == pytest.approx(2.0) and w[0] == pytest.approx((5 + 45) / 20)
    assert u[1] == pytest.approx(2.0)


def test_output_dataframe_and_sem():
    acc = SimpleNamespace(particle_counts=np.array([4, 1, 0]), particle_means=np.array([0.1, 0.2, 0.0]),
                          particle_m2s=np.array([0.12, 0.0, 0.0]), pair_sum_squares=np.array([2.0, 1.0, 0.0]),
                          pair_counts=np.array([8, 2, 0]))
    t_s = np.arange(10) * 2.0
    df = dmumu.build_output_dataframe([1, 2, 3], t_s, t_s * 0.5, acc)
    assert df["delta_mu2_particle_std"][0] == pytest.approx(math.sqrt(0.12 / 3))
    assert df["delta_mu2_particle_sem"][0] == pytest.approx(math.sqrt(0.12 / 3 / 4))
    assert math.isnan(df["delta_mu2_particle_sem"][1]) and math.isnan(df["delta_mu2_pair_mean"][2])
    assert df["tau_s"][2] == 6.0 and df["tau_gyroperiods"][0] == pytest.approx(1.0 / (2 * math.pi))


def test_dpp_arrays_and_snapshots():
    acc = dpp.new_accumulators(2)
    acc.counts[:] = [4, 0]
    acc.sum_delta_p[:] = [2.0, 0]
    acc.sum_delta_p2[:] = [2.0, 0]
    acc.sum_delta_p_norm[:] = [0.4, 0]
    acc.sum_delta_p_norm2[:] = [0.08, 0]
    out = dpp.compute_dpp_arrays(acc, [2.0, 4.0], [0.5, 1.0])
    assert out["dpp_centered_per_s"][0] == pytest.approx((0.5 - 0.25) / 4.0)
    assert out["dpp_raw_norm"][0] == pytest.approx(0.02 / 1.0)
    assert math.isnan(out["dpp_raw_per_s"][1])
    assert dpp.selected_energy_snapshot_indices(100, 5) == [1, 26, 50, 75, 100]  # 25.75 -> 26, 50.5 -> 50, 75.25 -> 75
    assert dpp.selected_energy_snapshot_indices(3, 16) == [1, 2, 3]


# --------------------------------------------------------------- generator

def test_generator_invariants():
    out = fields.generate_turbulence(17, B0_uG=5.0, dB_over_B0=0.5, u_rms_kms=20.0, kmin=1, kmax=5, seed=3,
                                     mode_decompose=True)
    Bx, By, Bz, ux, uy, uz = out["total"]
    assert Bx.shape == (17, 17, 17) and Bx.flags.f_contiguous  # Julia layout, x fastest
    # seam: the last node plane repeats the first on every axis
    for a in (Bx, uz):
        assert np.array_equal(a[0], a[-1]) and np.array_equal(a[:, 0], a[:, -1]) and np.array_equal(a[..., 0], a[..., -1])
    core = lambda a: a[:-1, :-1, :-1]  # noqa: E731
    dB = [core(Bx), core(By), core(Bz) - 5.0]
    assert math.sqrt(sum((d ** 2).mean() for d in dB)) == pytest.approx(2.5, rel=1e-12)
    du = [core(ux), core(uy), core(uz)]
    assert math.sqrt(sum((d ** 2).mean() for d in du)) == pytest.approx(20.0, rel=1e-12)
    # spectral divergence and band limits
    M = 16
    k = np.fft.fftfreq(M, 1.0 / M)
    Fs = [np.fft.fftn(d) for d in dB]  # axes of the (x,y,z) view
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    div = KX * Fs[0] + KY * Fs[1] + KZ * Fs[2]
    assert np.abs(div).max() <= 1e-9 * max(np.abs(f).max() for f in Fs)
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    power = sum(np.abs(f) ** 2 for f in Fs)
    assert power[(kk > 5) | (kk < 1)].max() <= 1e-20 * power.max()
    # modes sum to the total fluctuation (B and u); every magnetic part is
    # divergence-free
    for c in range(6):
        s = sum(core(out[m][c]) - (5.0 if c == 2 else 0.0) for m in ("alfven", "fast", "slow"))
        assert np.allclose(s, core(out["total"][c]) - (5.0 if c == 2 else 0.0), atol=1e-12)
    for m in ("alfven", "fast", "slow"):
        Fm = [np.fft.fftn(core(out[m][i]) - (5.0 if i == 2 else 0.0)) for i in range(3)]
        assert np.abs(KX * Fm[0] + KY * Fm[1] + KZ * Fm[2]).max() <= 1e-9 * max(np.abs(f).max() for f in Fs)
    # Alfven parts are transverse to B0 (no z component in B or u); the flow
    # of the fast and slow parts is orthogonal mode by mode
    Fa = [np.fft.fftn(core(out["alfven"][i]) - (5.0 if i == 2 else 0.0)) for i in range(6)]
    assert np.abs(Fa[2]).max() <= 1e-9 * np.abs(Fa[0]).max() and np.abs(Fa[5]).max() <= 1e-9 * np.abs(Fa[3]).max()
    Ff = [np.fft.fftn(core(out["fast"][i])) for i in range(3, 6)]
    Fsl = [np.fft.fftn(core(out["slow"][i])) for i in range(3, 6)]
    dot = sum(Ff[i] * np.conj(Fsl[i]) for i in range(3))
    assert np.abs(dot).max() <= 1e-9 * max(np.abs(f).max() for f in Ff) * max(np.abs(f).max() for f in Fsl)


def _mhd_operator(k, alpha):
    """Linear ideal-MHD wave operator on the displacement for B0 along z (V_A =
    1, c_s^2 = alpha): omega^2 xi = K xi with K = alpha k k^T + (k.z)^2 I
    + k k^T - (k.z)(k z^T + z k^T) (force of the perturbed pressure and
    Lorentz force, textbook form)."""
    z = np.array([0.0, 0.0, 1.0])
    kz = k @ z
    return alpha * np.outer(k, k) + kz * kz * np.eye(3) + np.outer(k, k) - kz * (np.outer(k, z) + np.outer(z, k))


@pytest.mark.parametrize("alpha", [0.05, 5.0 / 6.0, 1.0, 4.0])
def test_mode_directions_are_mhd_eigenvectors(alpha):
    """The generator's Alfven / fast / slow directions are orthonormal
    eigenvectors of the MHD wave operator, with the fast eigenvalue (omega^2)
    the largest and the slow one the smallest, for random k and along the
    axes."""
    rng = np.random.default_rng(11)
    ks = list(rng.standard_normal((40, 3))) + [np.array(v, float) for v in ((0, 0, 1), (0, 0, -2), (1, 0, 0),
                                                                              (0, 3, 0), (1, 1, 0), (1, 0, 1e-3))]
    for k in ks:
        k = k / np.linalg.norm(k)
        (eA, ef, es) = [np.array([float(c) for c in d]) for d in fields.mhd_eigendirections(
            np.array(k[0]), np.array(k[1]), np.array(k[2]), alpha)]
        Q = np.stack([eA, ef, es])
        assert np.allclose(Q @ Q.T, np.eye(3), atol=1e-12)
        Kop = _mhd_operator(k, alpha)
        lam = [float(e @ Kop @ e) for e in (eA, ef, es)]
        for e, l in zip((eA, ef, es), lam):
            assert np.allclose(Kop @ e, l * e, atol=1e-12)
        assert lam[0] == pytest.approx(k[2] ** 2, abs=1e-12)  # Alfven: omega = k_par V_A
        assert lam[1] >= lam[0] - 1e-12 and lam[2] <= lam[0] + 1e-12


def test_slow_mode_carries_no_field_at_low_beta():
    """At beta -> 0 the slow wave is sound along B0: its displacement is
    along z and it bends no field line, so the slow magnetic part vanishes
    while its flow does not."""
    out = fields.generate_turbulence(17, kmin=1, kmax=5, seed=3, mode_decompose=True, mode_beta=0.0)
    core = lambda a: a[:-1, :-1, :-1]  # noqa: E731
    slow_b = max(np.abs(core(out["slow"][i]) - (5.0 if i == 2 else 0.0)).max() for i in range(3))
    fast_b = max(np.abs(core(out["fast"][i]) - (5.0 if i == 2 else 0.0)).max() for i in range(3))
    slow_u = max(np.abs(core(out["slow"][i])).max() for i in range(3, 6))
    assert slow_b <= 1e-12 * fast_b and slow_u > 1.0


def test_generator_spectrum_slope():
    out = fields.generate_turbulence(65, dB_over_B0=1.0, kmin=1, kmax=32, seed=5)["total"]
    M = 64
    F = [np.fft.fftn(a[:-1, :-1, :-1] - a[:-1, :-1, :-1].mean()) for a in out[:3]]
    k = np.fft.fftfreq(M, 1.0 / M)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    P = sum(np.abs(f) ** 2 for f in F)
    shells = np.arange(2, 30)
    E = np.array([P[(kk >= s - 0.5) & (kk < s + 0.5)].sum() for s in shells])
    slope = np.polyfit(np.log(shells), np.log(E), 1)[0]
    assert slope == pytest.approx(-5.0 / 3.0, abs=0.1)


# --------------------------------------------------------------- injection

def _axes():
    return fields.build_uniform_coords(10.0, 9, 9, 9)


def test_injection_isotropic_and_momentum_magnitude():
    cfg = dict(seed=1, n_particles=20000)
    g0 = injection.energy_to_gamma(10.0)
    v0 = injection.energy_to_speed(10.0)
    x1, x2, x3, p1, p2, p3 = injection.init_particles(cfg, _axes(), g0, v0)
    pm = np.sqrt(p1 ** 2 + p2 ** 2 + p3 ** 2)
    assert np.allclose(pm, g0 * injection.M_P * v0, rtol=1e-14)
    assert abs((p3 / pm).mean()) < 0.02 and abs((p3 / pm) ** 2).mean() == pytest.approx(1 / 3, abs=0.01)
    L = fields.box_length_m(10.0)
    assert x1.min() >= 0 and x1.max() < L


def test_injection_fixed_mu_and_fixed_position():
    B = [np.random.default_rng(0).normal(size=(9, 9, 9)) for _ in range(3)]
    B[2] += 3.0
    axes = _axes()
    cfg = dict(seed=2, n_particles=500, injection_mode="fixed-mu", injection_mu0=0.3)
    g0, v0 = injection.energy_to_gamma(1.0), injection.energy_to_speed(1.0)
    x1, x2, x3, p1, p2, p3 = injection.init_particles(cfg, axes, g0, v0, np.float64, *B)
    b = [injection.trilinear_sample_host(Bc, x1, x2, x3, axes) for Bc in B]
    bn = np.sqrt(sum(c * c for c in b))
    pm = np.sqrt(p1 ** 2 + p2 ** 2 + p3 ** 2)
    mu = (p1 * b[0] + p2 * b[1] + p3 * b[2]) / (pm * bn)
    assert np.allclose(mu, 0.3, atol=1e-12)
    cfg = dict(seed=2, n_particles=4, injection_position_mode="fixed", injection_position=(0.5, 0.25, 1.0))
    x1, x2, x3, *_ = injection.init_particles(cfg, axes, g0, v0)
    L = fields.box_length_m(10.0)
    assert np.all(x1 == 0.5 * L) and np.all(x2 == 0.25 * L) and np.all(x3 == L)
    with pytest.raises(PipelineError):
        injection.init_particles(dict(cfg, injection_position=(1.5, 0, 0)), axes, g0, v0)
    with pytest.raises(PipelineError):
        injection.init_particles(dict(seed=1, n_particles=2, injection_mode="fixed-mu", injection_mu0=1.5), axes, g0,
                                 v0, np.float64, *B)
