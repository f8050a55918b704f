"""GPU parity: libpt_hip.so kernels against the CPU oracle on the same seeded
inputs.  Bars: integer outputs bit-exact; FP64 trajectories within 1e-10
relative step for step over 1000 steps; FP32 paths within the tolerances
stated per test."""
import math

import numpy as np
import pytest

import oracle
from particle_transport_amd import dmumu, dpp, fields
from particle_transport_amd.injection import C_LIGHT, M_P, Q_E
from particle_transport_amd.runner import Ensemble, Field

pytestmark = pytest.mark.gpu

L_BOX = fields.box_length_m(10.0)


def make_case(n_nodes=33, n=512, energy_gev=10.0, seed=11, kmax=8, dB=0.5, steps_per_gp=100):
    raw = fields.generate_turbulence(n_nodes, seed=seed, kmax=kmax, dB_over_B0=dB)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    rng = np.random.default_rng(seed + 1)
    x = rng.random((3, n)) * L_BOX
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    gamma = energy_gev * 1e9 * 1.602176634e-19 / (M_P * C_LIGHT ** 2) + 1
    p = d * gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / gamma ** 2)
    B0 = oracle.mean_abs_b(F[0], F[1], F[2])
    dt = 2 * math.pi * gamma * M_P / (Q_E * B0) / steps_per_gp
    g = oracle.make_geom(n_nodes, n_nodes, n_nodes, L_BOX)
    return F, x, p, dt, g


def rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def test_mean_abs_b_known_answer(ctx):
    # runtests.jl:221-225: fill(3), fill(4), fill(0) over 2x2x2 -> 5
    b = [np.full((2, 2, 2), v) for v in (3.0, 4.0, 0.0)]
    assert abs(ctx.mean_abs_b(*b) - 5.0) < 1e-15


def test_mean_abs_b_matches_oracle(ctx):
    F, *_ = make_case(n_nodes=41)
    ref = oracle.mean_abs_b(F[0], F[1], F[2])
    assert abs(ctx.mean_abs_b(F[0], F[1], F[2]) - ref) <= 1e-15 * ref
    fld = Field(ctx, F, L_BOX, "bf16_brick")
    assert abs(fld.B0 - ref) <= 1e-15 * ref  # B0 comes from the FP64 source values, not the BF16 copy
    fld.close()


@pytest.mark.parametrize("periodic", [True, False])
def test_ref_f64_phase_space_step_for_step(ctx, periodic):
    """FP64 pusher vs oracle, every step saved, 1000 steps: <= 1e-10 relative."""
    F, x, p, dt, g = make_case(n=256, energy_gev=100.0 if not periodic else 10.0)
    steps = 1000
    fld = Field(ctx, F, L_BOX, "f64_planar")
    ens = Ensemble(ctx, fld, x, p, pusher="ref_f64", periodic=periodic)
    r = ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=1, out_kind="phase",
                    out_dtype=np.float64, steps_per_launch=97)
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, 1, periodic, 1)
    live = np.isfinite(o["pos"])
    assert np.array_equal(live, np.isfinite(r.out_a))
    assert np.max(rel_err(r.out_a[live], o["pos"][live])) <= 1e-10
    assert np.max(rel_err(r.out_b[live], o["mom"][live])) <= 1e-10
    assert np.array_equal(r.alive_counts, o["alive_counts"])  # integer: exact
    xs, ps, alive = ens.state()
    assert np.array_equal(alive, o["alive"].astype(bool))
    ens.close()
    fld.close()


def test_ref_f64_mu_cache_and_chunked_calls(ctx):
    """mu at cadence (reconstruct_mu_step_kernel!) over three chunked calls vs one oracle run."""
    F, x, p, dt, g = make_case(n=300)
    total, stride = 601, 20
    fld = Field(ctx, F, L_BOX, "f64_planar")
    ens = Ensemble(ctx, fld, x, p, pusher="ref_f64")
    outs, alive = [], []
    for first, cnt in ((1, 250), (251, 101), (352, 250)):
        r = ens.advance(dt=dt, first_step=first, n_steps=cnt, total_steps=total, stride=stride, out_kind="mu",
                        out_dtype=np.float64)
        outs.append(r.out_a)
        alive.append(r.alive_counts)
    mu = np.concatenate([o for o in outs if o is not None])
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, total, total, stride, True, 2)
    assert mu.shape == o["mu"].shape == (len(range(1, total + 1, stride)), 300)
    assert np.max(np.abs(mu - o["mu"])) <= 1e-10
    ens.close()
    fld.close()


@pytest.mark.parametrize("pusher,storage", [("plain_f32", "f32_planar"), ("fast", "f32_planar"),
                                            ("fast", "f32_brick")])
def test_fp32_paths_track_fp64_oracle(ctx, pusher, storage):
    """Split-position FP32 pushers vs the FP64 oracle over 1000 steps at
    10 GeV: positions within 1e-4 cells, momenta within 1e-4 relative
    (accumulated FP32 rounding, ~sqrt(N) * 6e-8 per step plus chaos)."""
    F, x, p, dt, g = make_case(n=256)
    steps = 1000
    fld = Field(ctx, F, L_BOX, storage)
    ens = Ensemble(ctx, fld, x, p, pusher=pusher)
    ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps + 1, stride=1, steps_per_launch=128)
    xs, ps, alive = ens.state()
    ens.close()
    fld.close()
    xo, po = x.copy(), p.copy()
    oracle.run(F, g, xo, po, dt, 1, steps, steps + 1, 1, True, 0)
    W = g.xmax
    d = (xs - xo + W / 2) % W - W / 2  # difference on the periodic torus
    assert np.max(np.abs(d) / g.dx) <= 1e-4
    assert np.max(np.abs(ps - po) / np.linalg.norm(po, axis=0)) <= 1e-4
    assert alive.all()


@pytest.mark.parametrize("periodic", [True, False])
def test_planar_and_brick_fp32_fields_give_identical_trajectories(ctx, periodic):
    """FP32 planar and FP32 brick storage hold the same values and the
    coefficient reload computes from them identically, so the production
    pusher's outputs must be bit-identical.  At 1e5 GeV particles cross cells
    every few pushes: almost every push of every wave runs the
    wave-cooperative reload with several lanes at once (n = 1000 also leaves
    a partial last wave whose idle lanes serve as helpers)."""
    F, x, p, dt, g = make_case(n=1000, energy_gev=1e5, steps_per_gp=50)
    outs = []
    for storage in ("f32_planar", "f32_brick"):
        fld = Field(ctx, F, L_BOX, storage)
        ens = Ensemble(ctx, fld, x, p, pusher="fast", periodic=periodic)
        r = ens.advance(dt=dt, first_step=1, n_steps=600, total_steps=600, stride=3, out_kind="phase",
                        steps_per_launch=128)
        outs.append((r.out_a, r.out_b, r.alive_counts))
        ens.close()
        fld.close()
    (pa, ma, ca), (pb, mb, cb) = outs
    assert np.array_equal(ca, cb)
    assert np.array_equal(pa, pb, equal_nan=True)
    assert np.array_equal(ma, mb, equal_nan=True)
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, 600, 600, 3, periodic, 1)
    live = np.isfinite(o["pos"]) & np.isfinite(pa)
    assert np.max(np.abs(pa[live] - o["pos"][live]) / g.dx) <= 1e-3  # FP32 path vs FP64 over 600 pushes
    if not periodic:
        assert ca[-1] < 1000


@pytest.mark.parametrize("storage,tol", [("bf16_brick", 2e-2), ("f16_brick", 2e-3)])
def test_16bit_fields_mu_close_to_fp64(ctx, storage, tol):
    """BF16/FP16 brick fields (per-component offset + scale) vs FP64 fields:
    mu after 200 steps within tol (BF16 keeps 8 mantissa bits of dB)."""
    F, x, p, dt, g = make_case(n=512)
    steps, stride = 201, 50
    fld = Field(ctx, F, L_BOX, storage)
    ens = Ensemble(ctx, fld, x, p, pusher="fast")
    r = ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=stride, out_kind="mu")
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, stride, True, 2)
    err = np.abs(r.out_a - o["mu"])
    assert err[0].max() <= 4 * tol  # first sample: field decode error only
    assert np.median(err[-1]) <= tol
    ens.close()
    fld.close()


def test_fast_open_boundary_deaths_exact(ctx):
    """Open boundary at high energy: particles leave; alive counts per step and
    the NaN pattern of the mu cache equal the oracle's (integer work)."""
    F, x, p, dt, g = make_case(n=512, energy_gev=1e5, steps_per_gp=50)
    steps, stride = 400, 10
    # keep particles away from exact cell-boundary ties: compare where the
    # FP32 and FP64 exit steps agree (they differ only for grazing exits)
    fld = Field(ctx, F, L_BOX, "f32_brick")
    ens = Ensemble(ctx, fld, x, p, pusher="fast", periodic=False)
    r = ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=stride, out_kind="mu")
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, stride, False, 2)
    assert o["alive_counts"][-1] < 512  # some particles left the box
    dead_gpu = ~np.isfinite(r.out_a)
    dead_ref = ~np.isfinite(o["mu"])
    mismatch = np.sum(dead_gpu != dead_ref)
    assert mismatch <= 0.002 * dead_ref.size
    assert np.max(np.abs(r.alive_counts - o["alive_counts"])) <= 2
    ens.close()
    fld.close()


@pytest.mark.parametrize("pusher,storage", [("fast", "f32_brick"), ("fast", "f16_brick"),
                                            ("plain_f32", "f32_planar")])
def test_particles_on_the_upper_faces(ctx, pusher, storage):
    """x == L (or y, z, or all three) is inside the box and gathers the last
    cell at t = 1 (trilinear_sample's clamp of i0 to nx-1).  The production
    pusher evaluates that cell's polynomial with the offset folded into its
    coefficients: mu of the initial state must match the oracle there, and
    particles moving outward must die at the first push exactly as in the
    oracle (open boundary).  After that only particles the oracle leaves
    within FP32 offset resolution of a face (1e-7 cells) could differ; at
    1e4 GeV inward movers advance far more than that per push, so the counts
    and the NaN pattern must be exact."""
    F, x, p, dt, g = make_case(n=256, energy_gev=1e4)
    x[0, 0:64] = g.xmax
    x[1, 64:128] = g.ymax
    x[2, 128:192] = g.zmax
    x[:, 192:224] = np.array([[g.xmax], [g.ymax], [g.zmax]])
    fld = Field(ctx, F, L_BOX, storage)
    ens = Ensemble(ctx, fld, x, p, pusher=pusher, periodic=False)
    r = ens.advance(dt=dt, first_step=1, n_steps=3, total_steps=3, stride=1, out_kind="mu")
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, 3, 3, 1, False, 2)
    ph = oracle.run(F, g, x.copy(), p.copy(), dt, 1, 3, 3, 1, False, 1)
    tol = 2e-5 if storage != "f16_brick" else 2e-3
    assert np.max(np.abs(r.out_a[0] - o["mu"][0])) <= tol  # every particle alive at sample 1
    assert r.alive_counts[0] == o["alive_counts"][0]  # outward movers die at the first push
    assert o["alive_counts"][0] < 256 - 40
    faces = np.array([g.xmax, g.ymax, g.zmax])[None, :, None]
    with np.errstate(invalid="ignore"):
        gap = np.fmin.reduce((faces - ph["pos"][1:]) / g.dx, axis=1)  # cells inside the faces after each push
    grazer = np.nan_to_num(gap, nan=np.inf) < 1e-7
    assert np.max(np.abs(r.alive_counts - o["alive_counts"])) <= grazer.any(axis=0).sum()
    keep = ~grazer.any(axis=0)
    assert np.array_equal(np.isnan(r.out_a[:, keep]), np.isnan(o["mu"][:, keep]))
    live = ~np.isnan(o["mu"]) & ~np.isnan(r.out_a)
    assert np.max(np.abs(r.out_a[live] - o["mu"][live])) <= tol
    ens.close()
    fld.close()


def test_ref_f64_open_boundary_alive_counts_exact(ctx):
    F, x, p, dt, g = make_case(n=512, energy_gev=1e5, steps_per_gp=50)
    steps = 400
    fld = Field(ctx, F, L_BOX, "f64_planar")
    ens = Ensemble(ctx, fld, x, p, pusher="ref_f64", periodic=False)
    r = ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=10, out_kind="mu",
                    out_dtype=np.float64)
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, 10, False, 2)
    assert o["alive_counts"][-1] < 512
    assert np.array_equal(r.alive_counts, o["alive_counts"])
    assert np.array_equal(np.isfinite(r.out_a), np.isfinite(o["mu"]))
    ens.close()
    fld.close()


def _mu_series(n, nsave, seed, nan_frac=0.02):
    rng = np.random.default_rng(seed)
    mu = np.clip(np.cumsum(rng.normal(0, 0.05, (n, nsave)), axis=1) + rng.uniform(-1, 1, (n, 1)), -1, 1)
    mu = mu.astype(np.float32)
    mu[rng.random(mu.shape) < nan_frac] = np.nan
    mu[0, :] = 1.0  # right-closed last bin
    mu[1, ::3] = 0.5  # exactly on an edge
    return mu


@pytest.mark.parametrize("injection", [False, True])
@pytest.mark.parametrize("bin_abs", [True, False])
def test_dmumu_matches_oracle(ctx, injection, bin_abs):
    mu = _mu_series(700, 300, 5)
    lags = [1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 299, 300, 400]
    edges = np.linspace(0.0, 1.0, 21) if bin_abs else np.linspace(-1.0, 1.0, 41)
    nb = len(edges) - 1
    acc = dmumu.new_accumulators(len(lags), nb)
    oacc = oracle.new_dmumu_accumulators(len(lags), nb)
    for c0 in (0, 256):  # two chunks merged in place, as the reference does
        c1 = 256 if c0 == 0 else 700
        dmumu.process_mu_chunk_dmumu(ctx, mu[c0:c1], lags, edges, acc, "injection" if injection else "sliding",
                                     bin_abs)
        oracle.dmumu(mu[c0:c1], lags, edges, injection, bin_abs, oacc)
    assert np.array_equal(acc.dmumu_counts, oacc["dmumu_counts"])
    assert np.array_equal(acc.pair_counts, oacc["pair_counts"])
    assert np.array_equal(acc.particle_counts, oacc["particle_counts"])
    for k in ("dmumu_sum_delta", "dmumu_sum_delta2", "pair_sum_squares", "particle_means", "particle_m2s"):
        a, b = getattr(acc, k), oacc[k]
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12 * np.max(np.abs(b))), k


@pytest.mark.parametrize("injection", [False, True])
def test_dmumu_finite_walk_matches_oracle(ctx, injection):
    """Mostly finite series, as a pusher writes them: most particles finite
    throughout, some dead from a random sample on (NaN suffix), a few with
    isolated NaN samples, mixed within each wavefront.  Host time-fastest
    chunks, and a device particle-fastest cache with and without a selection
    list, all against the oracle."""
    from particle_transport_amd._lib import DeviceBuffer
    rng = np.random.default_rng(21)
    n, nsave = 700, 400
    mu = _mu_series(n, nsave, 22, nan_frac=0.0)
    dead = rng.random(n) < 0.1
    for i in np.nonzero(dead)[0]:
        mu[i, rng.integers(1, nsave):] = np.nan
    gaps = (rng.random(n) < 0.05) & ~dead
    for i in np.nonzero(gaps)[0]:
        mu[i, rng.integers(0, nsave, 3)] = np.nan
    lags = [1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 377, 399, 400]
    edges = np.linspace(0.0, 1.0, 21)
    mode = "injection" if injection else "sliding"

    def check(acc, oacc):
        assert np.array_equal(acc.dmumu_counts, oacc["dmumu_counts"])
        assert np.array_equal(acc.pair_counts, oacc["pair_counts"])
        assert np.array_equal(acc.particle_counts, oacc["particle_counts"])
        for k in ("dmumu_sum_delta", "dmumu_sum_delta2", "pair_sum_squares", "particle_means", "particle_m2s"):
            a, b = getattr(acc, k), oacc[k]
            assert np.allclose(a, b, rtol=1e-11, atol=1e-12 * np.max(np.abs(b))), k

    acc = dmumu.new_accumulators(len(lags), 20)
    dmumu.process_mu_chunk_dmumu(ctx, mu, lags, edges, acc, mode, True)
    oacc = oracle.new_dmumu_accumulators(len(lags), 20)
    oracle.dmumu(mu, lags, edges, injection, True, oacc)
    check(acc, oacc)
    dev = DeviceBuffer.from_host(ctx, np.ascontiguousarray(mu.T))
    acc = dmumu.new_accumulators(len(lags), 20)
    dmumu.process_mu_chunk_dmumu(ctx, dev, lags, edges, acc, mode, True, layout="particle_fastest")
    check(acc, oacc)
    sel = np.sort(rng.choice(n, 300, replace=False))
    acc = dmumu.new_accumulators(len(lags), 20)
    dmumu.process_mu_chunk_dmumu(ctx, dev, lags, edges, acc, mode, True, layout="particle_fastest",
                                 particle_index=sel)
    oacc = oracle.new_dmumu_accumulators(len(lags), 20)
    oracle.dmumu(mu[sel], lags, edges, injection, True, oacc)
    check(acc, oacc)
    dev.free()


def test_dmumu_device_cache_layout_and_selection(ctx):
    """mu cache [save][n] on the device with a particle selection list equals the
    reference's time-fastest chunk of the same particles."""
    from particle_transport_amd._lib import DeviceBuffer
    mu = _mu_series(1000, 200, 9)
    sel = np.sort(np.random.default_rng(2).choice(1000, 333, replace=False))
    lags = [1, 4, 16, 64, 150]
    edges = np.linspace(0.0, 1.0, 25)
    dev = DeviceBuffer.from_host(ctx, np.ascontiguousarray(mu.T))
    acc = dmumu.new_accumulators(len(lags), 24)
    dmumu.process_mu_chunk_dmumu(ctx, dev, lags, edges, acc, "sliding", True, layout="particle_fastest",
                                 particle_index=sel)
    oacc = oracle.new_dmumu_accumulators(len(lags), 24)
    oracle.dmumu(mu[sel], lags, edges, False, True, oacc)
    assert np.array_equal(acc.dmumu_counts, oacc["dmumu_counts"])
    assert np.allclose(acc.dmumu_sum_delta2, oacc["dmumu_sum_delta2"], rtol=1e-11)
    dev.free()


def test_dpp_and_energy_snapshots_match_oracle(ctx):
    rng = np.random.default_rng(4)
    nsave, n = 120, 600
    mom = (rng.normal(0, 1e-18, (nsave, 3, n)) + 5e-17).astype(np.float32)
    mom[rng.random((nsave, 3, n)) < 0.01] = np.nan
    lags = [1, 2, 7, 30, 119, 200]
    p0 = dpp.momentum_magnitude(*mom[0, :, 0])
    acc = dpp.new_accumulators(len(lags))
    dpp.process_momenta_chunk_dpp(ctx, mom, lags, p0, acc)
    oacc = oracle.new_dpp_accumulators(len(lags))
    oracle.dpp(mom, lags, p0, oacc)
    assert np.array_equal(acc.counts, oacc["counts"])
    for k in ("sum_delta_p", "sum_delta_p2", "sum_delta_p_norm", "sum_delta_p_norm2"):
        assert np.allclose(getattr(acc, k), oacc[k], rtol=1e-11, atol=1e-12 * np.max(np.abs(oacc[k]))), k
    snaps = dpp.selected_energy_snapshot_indices(nsave, 16)
    e = np.asfortranarray(np.full((len(snaps), n), np.nan))
    dpp.fill_energy_snapshots(ctx, e, mom, snaps, 1)
    eo = np.asfortranarray(np.full((len(snaps), n), np.nan))
    oracle.energy_snapshots(mom, np.asarray(snaps) - 1, eo, 0)
    assert np.array_equal(np.isnan(e), np.isnan(eo))
    assert np.allclose(e[np.isfinite(eo)], eo[np.isfinite(eo)], rtol=1e-14)


def test_dpp_finite_walk_matches_oracle(ctx):
    """Mostly finite momenta with some particles dead from a random sample on
    (NaN suffix), mixed within wavefronts, with and without a selection list."""
    rng = np.random.default_rng(12)
    nsave, n = 300, 900
    mom = (rng.normal(0, 1e-18, (nsave, 3, n)) + 5e-17).astype(np.float32)
    for i in np.nonzero(rng.random(n) < 0.1)[0]:
        mom[rng.integers(1, nsave):, :, i] = np.nan
    lags = [1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 299, 300]
    p0 = dpp.momentum_magnitude(*mom[0, :, 0])
    for sel in (None, np.sort(rng.choice(n, 400, replace=False))):
        acc = dpp.new_accumulators(len(lags))
        sub = mom if sel is None else mom[:, :, sel]
        dpp.process_momenta_chunk_dpp(ctx, mom, lags, p0, acc, particle_index=sel)
        oacc = oracle.new_dpp_accumulators(len(lags))
        oracle.dpp(sub, lags, p0, oacc)
        assert np.array_equal(acc.counts, oacc["counts"])
        for k in ("sum_delta_p", "sum_delta_p2", "sum_delta_p_norm", "sum_delta_p_norm2"):
            assert np.allclose(getattr(acc, k), oacc[k], rtol=1e-11, atol=1e-12 * np.max(np.abs(oacc[k]))), k


def test_dpp_from_fp64_pmag_cache_matches_oracle(ctx):
    """pt_dpp_accumulate_pmag (the serving path's FP64 |p| cache [save][n])
    equals the oracle's D_pp on the FP32 momenta those |p| come from."""
    from particle_transport_amd import _lib, serving
    rng = np.random.default_rng(7)
    nsave, n = 333, 1000
    mom = (rng.normal(0, 1e-18, (nsave, 3, n)) + 5e-17).astype(np.float32)
    mom[:, :, 5:9][rng.random((nsave, 3, 4)) < 0.05] = np.nan
    pm = np.sqrt((mom.astype(np.float64) ** 2).sum(axis=1))  # momentum_magnitude in FP64
    lags = list(range(1, 40, 3)) + [100, 332, 333]
    p0 = float(pm[0, 0])
    buf = _lib.DeviceBuffer.from_host(ctx, np.ascontiguousarray(pm))
    acc = serving.dpp_on_device(ctx, buf.ptr, nsave, n, lags, p0)
    buf.free()
    oacc = oracle.new_dpp_accumulators(len(lags))
    oracle.dpp(mom, lags, p0, oacc)
    assert np.array_equal(acc.counts, oacc["counts"])
    for k in ("sum_delta_p", "sum_delta_p2", "sum_delta_p_norm", "sum_delta_p_norm2"):
        assert np.allclose(getattr(acc, k), oacc[k], rtol=1e-11, atol=1e-12 * np.max(np.abs(oacc[k]))), k


@pytest.mark.parametrize("nsave,n", [(40, 40_000), (700, 300), (2600, 70), (6000, 40), (9000, 9)])
def test_dpp_tile_sizes_match_oracle(ctx, nsave, n):
    """The D_pp walk stages each particle's whole series in LDS, 16, 4 or 1
    particles per workgroup tile as the series length allows (40,000 particles
    of 40 samples: workgroups that take several tiles; 9000 FP64 samples: too
    long for a tile, the global walk).  Dead particles (NaN suffix) put some
    tiles on the checked path; lags reach past the series.  FP64 |p| from FP32
    momenta against the oracle (and bitwise equal over two calls), and the
    float32 relative series against an FP64 brute force."""
    from particle_transport_amd import _lib, serving
    rng = np.random.default_rng(nsave)
    mom = (rng.normal(0, 1e-18, (nsave, 3, n)) + 5e-17).astype(np.float32)
    for i in np.nonzero(rng.random(n) < 0.05)[0]:
        mom[rng.integers(1, nsave):, :, i] = np.nan
    mom[:, :, 0] = mom[0, :, 0]  # keep particle 0 alive: p0 below
    lags = sorted({1, 2, 3, nsave - 1, nsave, nsave + 7} | set(np.geomspace(4, nsave - 2, 8).astype(int).tolist()))
    p0 = dpp.momentum_magnitude(*mom[0, :, 0])
    acc = dpp.new_accumulators(len(lags))
    dpp.process_momenta_chunk_dpp(ctx, mom, lags, p0, acc)
    again = dpp.new_accumulators(len(lags))
    dpp.process_momenta_chunk_dpp(ctx, mom, lags, p0, again)
    for k in ("counts", "sum_delta_p", "sum_delta_p2", "sum_delta_p_norm", "sum_delta_p_norm2"):
        assert np.array_equal(getattr(acc, k), getattr(again, k)), k  # fixed merge order: bitwise run to run
    oacc = oracle.new_dpp_accumulators(len(lags))
    oracle.dpp(mom, lags, p0, oacc)
    assert np.array_equal(acc.counts, oacc["counts"]) and oacc["counts"].sum() > 0
    for k in ("sum_delta_p", "sum_delta_p2", "sum_delta_p_norm", "sum_delta_p_norm2"):
        assert np.allclose(getattr(acc, k), oacc[k], rtol=1e-11, atol=1e-12 * np.max(np.abs(oacc[k]))), k
    pm = np.sqrt((mom.astype(np.float64) ** 2).sum(axis=1))
    p_ref = float(np.nanmedian(pm))
    delta = np.ascontiguousarray((pm / p_ref - 1.0).astype(np.float32))
    buf = _lib.DeviceBuffer.from_host(ctx, delta)
    try:
        racc = serving.dpp_dprel_on_device(ctx, buf.ptr, nsave, n, lags, p_ref, p0)
    finally:
        buf.free()
    for li, lag in enumerate(lags):
        if lag >= nsave:
            assert racc.counts[li] == 0 and racc.sum_delta_p2[li] == 0.0
            continue
        a, b = delta[:-lag].astype(np.float64), delta[lag:].astype(np.float64)
        ok = np.isfinite(a) & np.isfinite(b)
        d = (b - a)[ok]
        assert racc.counts[li] == ok.sum()
        for k, want in (("sum_delta_p", p_ref * d.sum()), ("sum_delta_p2", p_ref ** 2 * (d * d).sum())):
            got = getattr(racc, k)[li]
            assert abs(got - want) <= 1e-11 * abs(want) + 1e-13 * abs(p_ref ** 2 * (d * d).sum()), (k, lag)


def test_dpp_from_relative_cache(ctx):
    """pt_dpp_accumulate_dprel over a float32 relative cache delta = |p|/p_ref - 1
    (the serving path's PT_OUT_DPREL): counts and sums equal a float64 brute
    force over the same delta series, Delta p = p_ref (delta(s+l) - delta(s)),
    and agree with the oracle's D_pp on the FP32 momenta the series was made
    from to the rounding of delta to float32."""
    from particle_transport_amd import _lib, serving
    rng = np.random.default_rng(11)
    nsave, n = 300, 700
    mom = (rng.normal(0, 1e-18, (nsave, 3, n)) + 5e-17).astype(np.float32)
    mom[:, :, 5:9][rng.random((nsave, 3, 4)) < 0.05] = np.nan
    pm = np.sqrt((mom.astype(np.float64) ** 2).sum(axis=1))
    p_ref, p0 = float(np.nanmedian(pm)), float(pm[0, 0])
    delta = np.ascontiguousarray((pm / p_ref - 1.0).astype(np.float32))
    lags = list(range(1, 40, 3)) + [100, 299, 300, 450]
    buf = _lib.DeviceBuffer.from_host(ctx, delta)
    acc = serving.dpp_dprel_on_device(ctx, buf.ptr, nsave, n, lags, p_ref, p0)
    buf.free()
    for li, lag in enumerate(lags):
        if lag >= nsave:  # the reference skips lags beyond the series
            assert acc.counts[li] == 0 and acc.sum_delta_p2[li] == 0.0
            continue
        a, b = delta[:-lag].astype(np.float64), delta[lag:].astype(np.float64)
        ok = np.isfinite(a) & np.isfinite(b)
        d = (b - a)[ok]
        assert acc.counts[li] == ok.sum()
        r = p_ref / p0
        for k, want in (("sum_delta_p", p_ref * d.sum()), ("sum_delta_p2", p_ref ** 2 * (d * d).sum()),
                        ("sum_delta_p_norm", r * d.sum()), ("sum_delta_p_norm2", r * r * (d * d).sum())):
            got = getattr(acc, k)[li]
            assert abs(got - want) <= 1e-11 * abs(want) + 1e-13 * abs(p_ref ** 2 * (d * d).sum()), (k, lag)
    oacc = oracle.new_dpp_accumulators(len(lags))
    oracle.dpp(mom, lags, p0, oacc)
    assert np.array_equal(acc.counts, oacc["counts"])
    for k in ("sum_delta_p2", "sum_delta_p_norm2"):
        assert np.allclose(getattr(acc, k), oacc[k], rtol=1e-6, atol=0), k


@pytest.mark.parametrize("storage", ["f64_planar", "f32_planar"])
@pytest.mark.parametrize("steps,stride", [(300, 30), (1200, 5)], ids=["10_samples", "240_samples"])
def test_mu_from_phase_matches_oracle(ctx, storage, steps, stride):
    """Both output layouts against the oracle's reconstruct_mu_chunk; the
    long case spans two 128-sample thread segments, and the particles move
    across cells within a segment (the cached corners are reloaded)."""
    F, x, p, dt, g = make_case(n=200, energy_gev=1e4)
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, stride, True, 1)
    pos, mom = o["pos"].astype(np.float32), o["mom"].astype(np.float32)
    pos[3, :, 7] = np.nan
    pos[5, 1, 9] = -1.0  # outside the box: NaN like the reference
    fld = Field(ctx, F, L_BOX, storage)
    from particle_transport_amd import _lib
    nsave, _, n = pos.shape
    out = np.empty((n, nsave), dtype=np.float32)
    _lib.check(_lib.load().pt_mu_from_phase(ctx.h, fld.h, n, nsave, _lib.PT_F32, 0, _lib.ptr(pos), _lib.ptr(mom),
                                            _lib.PT_F32, 0, _lib.ptr(out)))
    sm = np.empty((nsave, n), dtype=np.float32)
    _lib.check(_lib.load().pt_mu_from_phase_layout(ctx.h, fld.h, n, nsave, _lib.PT_F32, 0, _lib.ptr(pos),
                                                   _lib.ptr(mom), _lib.PT_F32, 0, _lib.ptr(sm), 1))
    assert np.array_equal(sm, out.T, equal_nan=True)
    ref = oracle.reconstruct_mu_chunk(pos.astype(np.float64), mom.astype(np.float64), F[:3], g)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    assert np.isnan(out[9, 5]) and np.isnan(out[7, 3])
    tol = 1e-12 if storage == "f64_planar" else 2e-5  # FP32: the reference's compute_precision=Float32
    assert np.nanmax(np.abs(out - ref)) <= tol + 1e-7  # + float32 output rounding
    if nsave > 64:
        cells = np.floor(o["pos"] / g.dx)
        assert np.any(np.diff(cells, axis=0) != 0)  # cell changes inside the series
    fld.close()


@pytest.mark.parametrize("pusher,storage", [("fast", "bf16_brick"), ("plain_f32", "f32_planar")])
def test_fp32_momentum_magnitude_conserved_in_uniform_B(ctx, pusher, storage):
    """Uniform B, u = 0: Boris conserves |p| exactly and turns p by 2*atan(|t|)
    per push.  With |p|^2 carried in FP64 the FP32 pusher keeps |p| to 2e-7
    over 20,000 pushes (uncorrected FP32 rotation drifts ~5e-8 per push) and
    the gyrophase to 1e-5 rad per gyration."""
    n_nodes, n, steps = 9, 256, 20000
    Bz = 5e-10
    F = [np.asfortranarray(np.full((n_nodes,) * 3, v)) for v in (0.0, 0.0, Bz, 0.0, 0.0, 0.0)]
    rng = np.random.default_rng(0)
    x = rng.random((3, n)) * L_BOX
    gamma = 50.0
    pm = M_P * C_LIGHT * math.sqrt(gamma ** 2 - 1)
    ang0 = rng.random(n) * 2 * math.pi
    pz = pm * rng.uniform(-0.9, 0.9, n)
    pperp = np.sqrt(pm ** 2 - pz ** 2)
    p = np.stack([pperp * np.cos(ang0), pperp * np.sin(ang0), pz])
    Omega = Q_E * Bz / (gamma * M_P)
    dt = 2 * math.pi / Omega / 200
    fld = Field(ctx, F, L_BOX, storage)
    ens = Ensemble(ctx, fld, x, p, pusher=pusher)
    n_run = 2000 if pusher == "plain_f32" else steps
    ens.advance(dt=dt, first_step=1, n_steps=n_run, total_steps=n_run + 1, stride=1, steps_per_launch=500)
    _, ps, _ = ens.state()
    ens.close()
    fld.close()
    mag = np.linalg.norm(ps, axis=0)
    assert np.max(np.abs(mag / pm - 1)) <= 2e-7
    t = Q_E * Bz * dt / (2 * M_P * gamma)
    expected = ang0 - n_run * 2 * math.atan(t)
    got = np.arctan2(ps[1], ps[0])
    dphi = (got - expected + math.pi) % (2 * math.pi) - math.pi
    assert np.max(np.abs(dphi)) <= 1e-5 * n_run / 200 + 1e-6
    assert np.allclose(ps[2], pz, rtol=1e-6)


@pytest.mark.parametrize("shape", [(9, 9, 9), (9, 17, 5)])
def test_rotation_angle_at_any_pushes_per_gyration(ctx, shape):
    """Uniform B, u = 0: every push turns p by 2*atan(|t|) with t = q B dt / (2 m gamma)
    and keeps |p|.  Lorentz factors from 1.2 to 60 at one dt give 4 to 200
    pushes per gyration, so lanes with t^2 on both sides of 2^-7 (where the
    pusher switches from the series for 2t/(1+t^2) to the reciprocal) share
    every wave.  The (9, 17, 5) grid has a different cell size on each axis
    (the general position step), the (9, 9, 9) grid cubic cells."""
    n, n_run = 2048, 2000
    Bz = 5e-10
    F = [np.asfortranarray(np.full(shape, v)) for v in (0.0, 0.0, Bz, 0.0, 0.0, 0.0)]
    rng = np.random.default_rng(4)
    x = rng.random((3, n)) * L_BOX
    gamma = rng.uniform(1.2, 60.0, n)
    pm = M_P * C_LIGHT * np.sqrt(gamma ** 2 - 1)
    ang0 = rng.random(n) * 2 * math.pi
    pz = pm * rng.uniform(-0.9, 0.9, n)
    pperp = np.sqrt(pm ** 2 - pz ** 2)
    p = np.stack([pperp * np.cos(ang0), pperp * np.sin(ang0), pz])
    dt = 2 * math.pi * 60.0 * M_P / (Q_E * Bz) / 200
    t = Q_E * Bz * dt / (2 * M_P * gamma)
    assert (t ** 2 > 2.0 ** -7).mean() > 0.1 and (t ** 2 < 2.0 ** -7).mean() > 0.5
    fld = Field(ctx, F, L_BOX, "f32_brick")
    ens = Ensemble(ctx, fld, x, p, pusher="fast")
    ens.advance(dt=dt, first_step=1, n_steps=n_run, total_steps=n_run + 1, stride=1, steps_per_launch=500)
    xs, ps, alive = ens.state()
    ens.close()
    fld.close()
    assert alive.all()
    assert np.max(np.abs(np.linalg.norm(ps, axis=0) / pm - 1)) <= 2e-7
    theta = 2 * np.arctan(t)
    expected = ang0 - n_run * theta
    dphi = (np.arctan2(ps[1], ps[0]) - expected + math.pi) % (2 * math.pi) - math.pi
    # FP32 t carries a relative error of a few 1e-7 (rsq, folded factors);
    # the series without its fallback would be off by 13 % at t^2 = 0.6
    assert np.all(np.abs(dphi) <= n_run * theta * 1e-6 + 1e-6)
    assert np.allclose(ps[2], pz, rtol=1e-6)
    # the guiding centre drifts along z at v_par, which the oracle fixes too
    xo, po = x.copy(), p.copy()
    g = oracle.make_geom(*shape, L_BOX)
    oracle.run(F, g, xo, po, dt, 1, n_run, n_run + 1, 1, True, 0)
    W = g.xmax
    d = (xs - xo + W / 2) % W - W / 2
    assert np.max(np.abs(d[2]) / g.dz) <= 1e-3


def test_fast_pusher_on_non_cubic_cells_tracks_oracle(ctx):
    """A turbulent field on a 33 x 17 x 25 grid (three cell sizes): the
    production pusher's general position step vs the FP64 oracle, to the
    same bounds as on cubic cells."""
    Fc, x, p, dt, _ = make_case(n=256)
    F = [np.asfortranarray(a[:, :17, :25]) for a in Fc]
    g = oracle.make_geom(33, 17, 25, L_BOX)
    steps = 1000
    fld = Field(ctx, F, L_BOX, "f32_brick")
    ens = Ensemble(ctx, fld, x, p, pusher="fast")
    ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps + 1, stride=1, steps_per_launch=128)
    xs, ps, alive = ens.state()
    ens.close()
    fld.close()
    xo, po = x.copy(), p.copy()
    oracle.run(F, g, xo, po, dt, 1, steps, steps + 1, 1, True, 0)
    W = np.array([[g.xmax], [g.ymax], [g.zmax]])
    d = (xs - xo + W / 2) % W - W / 2
    assert np.max(np.abs(d) / np.array([[g.dx], [g.dy], [g.dz]])) <= 1e-4
    assert np.max(np.abs(ps - po) / np.linalg.norm(po, axis=0)) <= 1e-4
    assert alive.all()


@pytest.mark.parametrize("periodic", [True, False])
def test_morton_sort_leaves_results_bit_identical(ctx, periodic):
    """Re-sorting changes which lane pushes which particle, never the
    arithmetic: mu caches, final state and alive counts are bit-identical."""
    F, x, p, dt, g = make_case(n=5000, energy_gev=1e4, steps_per_gp=50)
    fld = Field(ctx, F, L_BOX, "bf16_brick")
    runs = []
    for sort_every in (0, 150):
        ens = Ensemble(ctx, fld, x, p, pusher="fast", periodic=periodic)
        r = ens.advance(dt=dt, first_step=1, n_steps=600, total_steps=600, stride=20, out_kind="mu",
                        steps_per_launch=64, sort_every_steps=sort_every)
        runs.append((r, ens.state()))
        ens.close()
    fld.close()
    (ra, (xa, pa, aa)), (rb, (xb, pb, ab)) = runs
    assert np.array_equal(np.isnan(ra.out_a), np.isnan(rb.out_a))
    assert np.array_equal(ra.out_a[~np.isnan(ra.out_a)], rb.out_a[~np.isnan(rb.out_a)])
    assert np.array_equal(ra.alive_counts, rb.alive_counts)
    assert np.array_equal(aa, ab) and np.array_equal(xa[:, aa], xb[:, ab]) and np.array_equal(pa, pb)


def _dmumu_gpu(ctx, mu, lags, edges, injection, bin_abs):
    acc = dmumu.new_accumulators(len(lags), len(edges) - 1)
    dmumu.process_mu_chunk_dmumu(ctx, mu, lags, edges, acc, "injection" if injection else "sliding", bin_abs)
    return acc


def _check_dmumu(acc, oacc):
    assert np.array_equal(acc.dmumu_counts, oacc["dmumu_counts"])
    assert np.array_equal(acc.pair_counts, oacc["pair_counts"])
    assert np.array_equal(acc.particle_counts, oacc["particle_counts"])
    for k in ("dmumu_sum_delta", "dmumu_sum_delta2", "pair_sum_squares", "particle_means", "particle_m2s"):
        a, b = getattr(acc, k), oacc[k]
        scale = np.max(np.abs(b)) if b.size and np.max(np.abs(b)) > 0 else 1.0
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12 * scale), k


@pytest.mark.parametrize("case", ["one_particle_two_samples", "max_bins", "one_bin", "all_nan", "lags_past_end"])
def test_dmumu_edge_shapes(ctx, case):
    """Shapes at the edges the reference's loops allow: a single particle
    with two samples, 256 bins (the kernel's maximum) and one bin, particles
    whose series are all NaN, and lags at or past the series end (skipped)."""
    if case == "one_particle_two_samples":
        mu, lags, edges = np.array([[0.3, 0.7]], np.float32), [1, 2, 5], np.linspace(0.0, 1.0, 21)
    elif case == "max_bins":
        mu, lags, edges = _mu_series(300, 200, 3), [1, 7, 50, 199], np.linspace(0.0, 1.0, 257)
    elif case == "one_bin":
        mu, lags, edges = _mu_series(300, 200, 4), [1, 7, 50, 199], np.array([0.0, 1.0])
    elif case == "all_nan":
        mu, lags, edges = np.full((70, 50), np.nan, np.float32), [1, 2, 3], np.linspace(0.0, 1.0, 11)
    else:
        mu, lags, edges = _mu_series(130, 40, 5), [39, 40, 41, 1000], np.linspace(0.0, 1.0, 11)
    for injection in (False, True):
        for bin_abs in (True, False):
            e = edges if bin_abs else np.linspace(-1.0, 1.0, len(edges))
            acc = _dmumu_gpu(ctx, mu, lags, e, injection, bin_abs)
            oacc = oracle.new_dmumu_accumulators(len(lags), len(e) - 1)
            oracle.dmumu(mu, lags, e, injection, bin_abs, oacc)
            _check_dmumu(acc, oacc)
    if case == "all_nan":
        assert acc.dmumu_counts.sum() == 0 and acc.particle_counts.sum() == 0


def test_dmumu_offsets_beyond_32_bits(ctx):
    """A time-fastest cache of 2,097,664 particles x 512 samples (4.3 GB):
    particles selected near its end sit more than 2^32 bytes from its start,
    so the kernel takes its 64-bit addressing walk.  Only the selected
    particles' series are written; the oracle sees the same series."""
    import ctypes

    from particle_transport_amd import _lib
    from particle_transport_amd._lib import DeviceBuffer
    total, nsave = (1 << 21) + 512, 512
    assert (total - 1) * nsave * 4 > (1 << 32)
    rng = np.random.default_rng(8)
    sel = np.sort(np.concatenate([rng.choice(1000, 40, replace=False),
                                  total - 1 - rng.choice(2000, 60, replace=False)])).astype(np.int64)
    series = _mu_series(len(sel), nsave, 12, nan_frac=0.002)
    dev = DeviceBuffer(ctx, (total, nsave), np.float32)
    try:
        for k, j in enumerate(sel):
            row = np.ascontiguousarray(series[k])
            _lib.check(_lib.load().pt_memcpy(ctx.h, ctypes.c_void_p(dev.ptr + int(j) * nsave * 4),
                                             row.ctypes.data_as(ctypes.c_void_p), row.nbytes, 0))
        lags = [1, 3, 10, 64, 300, 511]
        edges = np.linspace(0.0, 1.0, 21)
        for injection in (False, True):
            acc = dmumu.new_accumulators(len(lags), 20)
            dmumu.process_mu_chunk_dmumu(ctx, dev, lags, edges, acc, "injection" if injection else "sliding", True,
                                         layout="time_fastest", particle_index=sel)
            oacc = oracle.new_dmumu_accumulators(len(lags), 20)
            oracle.dmumu(series, lags, edges, injection, True, oacc)
            _check_dmumu(acc, oacc)
    finally:
        dev.free()


# The motional field E = -u x B of every pusher (phase_space_gpu_runner.jl:365-368),
# checked as the oracle is (tests/test_oracle.py): in B = B0 z with a uniform
# flow u = u0 x, the guiding centre drifts at +u0 x and a particle at rest in
# the flow frame keeps v = u0 x.  A flipped sign or swapped operands drift at -u0.
@pytest.mark.parametrize("pusher,storage", [("ref_f64", "f64_planar"), ("plain_f32", "f32_planar"),
                                            ("fast", "f32_planar"), ("fast", "f32_brick"),
                                            ("fast", "f16_brick"), ("fast", "bf16_brick")])
def test_motional_field_drifts_with_the_flow(ctx, pusher, storage):
    B0, u0, gamma = 5e-10, 3.0e6, 1.0001
    n_nodes, Lb = 5, 1.0e12
    F = [np.full((n_nodes,) * 3, v, order="F") for v in (0.0, 0.0, B0, u0, 0.0, 0.0)]
    vmag = C_LIGHT * math.sqrt(1 - 1 / gamma ** 2)
    g1 = 1 / math.sqrt(1 - (u0 / C_LIGHT) ** 2)
    x = np.array([[0.4 * Lb, 0.4 * Lb], [0.5 * Lb, 0.5 * Lb], [0.5 * Lb, 0.5 * Lb]])
    p = np.array([[0.0, g1 * M_P * u0], [gamma * M_P * vmag, 0.0], [0.0, 0.0]])
    dt = 2 * math.pi * gamma * M_P / (Q_E * B0) / 100
    nsteps, stride = 4001, 10  # 40 gyrations
    fld = Field(ctx, F, Lb, storage)
    ens = Ensemble(ctx, fld, x, p, pusher=pusher, periodic=False)
    r = ens.advance(dt=dt, first_step=1, n_steps=nsteps, total_steps=nsteps, stride=stride, out_kind="phase",
                    out_dtype=np.float64, steps_per_launch=200)
    ens.close()
    fld.close()
    pos = r.out_a  # (nsave, 3, 2)
    assert np.isfinite(pos).all() and list(r.alive_counts[-1:]) == [2]
    t = np.arange(pos.shape[0]) * stride * dt
    assert np.polyfit(t, pos[:, 0, 0], 1)[0] == pytest.approx(u0, rel=0.02)
    assert abs(np.polyfit(t, pos[:, 1, 0], 1)[0]) <= 0.02 * u0
    # at rest in the flow frame: the straight line x0 + u0 t.  FP64 keeps it to
    # the Boris step's O((u0/c)^2) balance (1e-6 of the path).  The plain FP32
    # path adds the same increment u0 dt / dx = 1.6e-5 cells to an offset in
    # [0.5, 1) every push, and each sum rounds the same way, so its split
    # position can lag or lead by up to half an ulp (3e-8 cells) per push: at
    # most 0.5 * 2^-24 / 1.6e-5 = 1.9e-3 of the path.  The production pusher
    # adds each step with compensated summation, so its position error stays
    # at the FP32 velocity's resolution (1e-5 of the path allows 160 ulps of
    # the offset).  A flipped E instead gyrates with radius 1.6 % of the path.
    path = u0 * t[-1]
    tol = {"ref_f64": 1e-6, "plain_f32": 2e-3}.get(pusher, 1e-5)
    print(pusher, storage, "straight-line deviation / path:",
          float(np.max(np.abs(pos[:, 0, 1] - (0.4 * Lb + u0 * t))) / path))
    assert np.max(np.abs(pos[:, 0, 1] - (0.4 * Lb + u0 * t))) <= tol * path
    assert np.max(np.abs(pos[:, 1, 1] - 0.5 * Lb)) <= tol * path


def test_dmumu_bin_sums_deterministic_and_order_independent(ctx):
    """The per-bin sums are exact 128-bit fixed-point sums: the same call
    twice gives bitwise equal accumulators, and visiting the particles in
    another order (a permuted selection, so other lanes and workgroups flush
    into each bin in another order) leaves counts and bin sums bitwise
    unchanged.  (The reference merges per-thread accumulators in a fixed
    order, compute_delta_mu2_dmumu_full.jl:540-557.)  The oracle pins the
    values."""
    mu = _mu_series(20000, 400, 21, nan_frac=0.001)  # time fastest: (particles, samples)
    lags = [1, 2, 4, 8, 16, 32, 64, 128, 256, 399]
    edges = np.linspace(0.0, 1.0, 41)
    runs = []
    for perm_seed in (None, None, 3):
        idx = None if perm_seed is None else np.random.default_rng(perm_seed).permutation(len(mu)).astype(np.int64)
        acc = dmumu.new_accumulators(len(lags), 40)
        dmumu.process_mu_chunk_dmumu(ctx, mu, lags, edges, acc, "sliding", True, particle_index=idx)
        runs.append(acc)
    for f in ("dmumu_counts", "dmumu_sum_delta", "dmumu_sum_delta2", "particle_counts", "particle_means",
              "particle_m2s", "pair_sum_squares", "pair_counts"):
        assert np.array_equal(getattr(runs[0], f), getattr(runs[1], f)), f
    for f in ("dmumu_counts", "dmumu_sum_delta", "dmumu_sum_delta2", "pair_counts"):
        assert np.array_equal(getattr(runs[0], f), getattr(runs[2], f)), f
    oacc = oracle.new_dmumu_accumulators(len(lags), 40)
    oracle.dmumu(mu, lags, edges, False, True, oacc)
    _check_dmumu(runs[0], oacc)


def _fixed_point_bins(mu, lags, edges, bin_abs):
    """The bin sums as the kernel defines them, in exact integer arithmetic:
    each particle's starts are split into runs of one start bin; a run's sums
    of dmu and dmu^2 over each lag are formed in FP64 in start order (dmu^2
    added with one rounding, as a fused multiply-add); each run sum is
    truncated toward zero to a multiple of 2^-80; the bin's total is the exact
    sum of those integers modulo 2^128, read back as hi 2^-16 + lo 2^-80 in
    FP64.  Runs whose start lies outside the edges are dropped."""
    from fractions import Fraction
    nl, nb = len(lags), len(edges) - 1
    cnt = np.zeros((nl, nb), np.int64)
    tot = [[[0, 0] for _ in range(nb)] for _ in range(nl)]

    def key(a):
        v = abs(float(a)) if bin_abs else float(a)
        if v < edges[0]:
            return -1
        if v > edges[-1]:
            return -2
        return min(int(np.searchsorted(edges, v, side="right")) - 1, nb - 1)

    def fix(v):
        q = Fraction(v) * 2 ** 80
        return int(q) if q >= 0 else -int(-q)

    nsave = mu.shape[1]
    for row in mu:
        for g, lag in enumerate(lags):
            cur, rc, r1, r2 = None, 0, 0.0, 0.0

            def flush():
                if cur is not None and cur >= 0 and rc:
                    cnt[g, cur] += rc
                    tot[g][cur][0] += fix(r1)
                    tot[g][cur][1] += fix(r2)

            for s in range(nsave - 1):
                a = row[s]
                if not np.isfinite(a):
                    continue
                k = key(a)
                if k != cur:
                    flush()
                    cur, rc, r1, r2 = k, 0, 0.0, 0.0
                if s + lag < nsave and np.isfinite(row[s + lag]):
                    d = float(row[s + lag]) - float(a)
                    r1 += d
                    r2 = float(Fraction(d) * Fraction(d) + Fraction(r2))
                    rc += 1
            flush()

    def to_double(t):
        t %= 1 << 128
        lo, hi = t & ((1 << 64) - 1), t >> 64
        if hi >= 1 << 63:
            hi -= 1 << 64
        return float(hi) * 2.0 ** -16 + float(lo) * 2.0 ** -80

    s1 = np.array([[to_double(tot[g][b][0]) for b in range(nb)] for g in range(nl)])
    s2 = np.array([[to_double(tot[g][b][1]) for b in range(nb)] for g in range(nl)])
    return cnt, s1, s2


def test_dmumu_bin_sums_equal_exact_fixed_point_sums(ctx):
    """Bitwise equality of the bin sums with an exact-integer restatement of
    their fixed-point definition: pins the conversion of run sums to 2^-80
    units (truncation of both signs, zeros), the carry-free LDS words and the
    128-bit carries across workgroups (three tiles of 256 particles).  Signed
    mu, NaN samples, runs below and above the edges, constant series (long
    runs) and lags reaching the end of the series."""
    rng = np.random.default_rng(77)
    n, nsave = 600, 90
    mu = np.clip(np.cumsum(rng.normal(0, 0.08, (n, nsave)), axis=1) + rng.uniform(-1, 1, (n, 1)), -1, 1)
    mu = mu.astype(np.float32)
    mu[rng.random(mu.shape) < 0.01] = np.nan
    mu[5, :] = np.float32(0.3)
    mu[300, :] = -np.float32(0.7)
    mu[301, ::2] = np.float32(0.25)
    lags = [1, 2, 5, 17, 89]
    for bin_abs, edges in ((True, np.linspace(0.1, 0.9, 7)), (False, np.linspace(-0.8, 0.95, 9))):
        acc = dmumu.new_accumulators(len(lags), len(edges) - 1)
        dmumu.process_mu_chunk_dmumu(ctx, mu, lags, edges, acc, "sliding", bin_abs)
        cnt, s1, s2 = _fixed_point_bins(mu, lags, edges, bin_abs)
        assert np.array_equal(acc.dmumu_counts, cnt)
        assert np.array_equal(acc.dmumu_sum_delta, s1)
        assert np.array_equal(acc.dmumu_sum_delta2, s2)
        assert (s1 < 0).any() and (s1 > 0).any()


@pytest.mark.parametrize("pusher,storage", [("fast", "f32_brick"), ("plain_f32", "f32_planar")])
def test_periodic_particles_on_the_upper_faces_wrap(ctx, pusher, storage):
    """Periodic box, particles injected exactly on x == L (or y, z, or all
    three), half moving outward: the reference samples the field there
    (trilinear_sample's clamp) and wraps x > L to x - L after the push
    (wrap_periodic_coordinate, phase_space_gpu_runner.jl:325-328,372-376).
    Every cached position must stay in [0, L] and match the oracle on the
    torus; mu at the first sample must match at the face."""
    F, x, p, dt, g = make_case(n=256, energy_gev=1e4)
    x[0, 0:64] = g.xmax
    x[1, 64:128] = g.ymax
    x[2, 128:192] = g.zmax
    x[:, 192:224] = np.array([[g.xmax], [g.ymax], [g.zmax]])
    fld = Field(ctx, F, L_BOX, storage)
    ens = Ensemble(ctx, fld, x, p, pusher=pusher, periodic=True)
    steps = 20
    r = ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=1, out_kind="phase",
                    out_dtype=np.float64)
    ens.close()
    o = oracle.run(F, g, x.copy(), p.copy(), dt, 1, steps, steps, 1, True, 1)
    pos = r.out_a
    assert np.isfinite(pos).all()
    assert pos.min() >= 0.0 and pos.max() <= g.xmax
    W = g.xmax
    d = (pos - o["pos"] + W / 2) % W - W / 2
    assert np.max(np.abs(d) / g.dx) <= 1e-4
    assert np.max(np.abs(r.out_b - o["mom"]) / np.linalg.norm(o["mom"], axis=1, keepdims=True)) <= 1e-4
    out_mu = Ensemble(ctx, fld, x, p, pusher=pusher, periodic=True)
    m = out_mu.advance(dt=dt, first_step=1, n_steps=1, total_steps=2, stride=1, out_kind="mu")
    out_mu.close()
    fld.close()
    om = oracle.run(F, g, x.copy(), p.copy(), dt, 1, 1, 2, 1, True, 2)
    assert np.max(np.abs(m.out_a[0] - om["mu"][0])) <= 2e-5
