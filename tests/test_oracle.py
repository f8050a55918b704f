"""Pin the CPU oracle before trusting it: each check is independent of the
oracle's own formulas (closed forms, invariants, brute force in plain Python
on tiny inputs), chosen so a dropped term, wrong sign/index or transposed
operand fails one of them."""
import math

import numpy as np
import pytest

import oracle
from particle_transport_amd import dmumu as host_dmumu
from particle_transport_amd.injection import C_LIGHT, M_P, Q_E

L = 3.0e17


def uniform_fields(n, B=(0.0, 0.0, 5e-10), u=(0.0, 0.0, 0.0)):
    return [np.full((n, n, n), v, order="F") for v in (*B, *u)]


# ---------------------------------------------------------------- trilinear

def test_trilinear_exact_on_trilinear_polynomial():
    n = 9
    g = oracle.make_geom(n, n, n, L)
    xs = np.arange(n) * g.dx
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    f = lambda x, y, z: 1.5 + 2.0 * x / L - 3.0 * y / L + 0.5 * z / L + 4.0 * x * y / L ** 2 \
        - 2.5 * x * z / L ** 2 + 1.25 * y * z / L ** 2 + 7.0 * x * y * z / L ** 3  # noqa: E731
    arr = np.asfortranarray(f(X, Y, Z))
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rng.random(3) * L
        assert abs(oracle.trilinear(arr, *p, g) - f(*p)) <= 1e-12 * 10


def test_trilinear_at_nodes_and_clamped_edge():
    n = 5
    g = oracle.make_geom(n, n, n, L)
    arr = np.asfortranarray(np.random.default_rng(1).random((n, n, n)))
    for i, j, k in [(0, 0, 0), (1, 2, 3), (4, 4, 4), (3, 0, 4)]:
        assert oracle.trilinear(arr, i * g.dx, j * g.dx, k * g.dx, g) == pytest.approx(arr[i, j, k], abs=1e-14)
    # x == L gathers the last cell at t = 1 (clamp of i0 to nx-1)
    assert oracle.trilinear(arr, L, 0.0, 0.0, g) == pytest.approx(arr[4, 0, 0], abs=1e-14)


# -------------------------------------------------------------------- Boris

def test_boris_uniform_B_rotation_angle_and_norm():
    Bz = 5e-10
    gamma = 10.0
    pmag = M_P * C_LIGHT * math.sqrt(gamma ** 2 - 1)
    p = np.array([pmag, 0.0, 0.0])
    Omega = Q_E * Bz / (gamma * M_P)
    dt = 0.05 / Omega
    q, _ = oracle.boris_push(p, (0, 0, 0), (0, 0, 0), (0, 0, Bz), dt)
    assert np.linalg.norm(q) == pytest.approx(pmag, rel=1e-15)
    t = Q_E * Bz * dt / (2 * M_P * gamma)
    ang = math.atan2(q[1], q[0])
    assert ang == pytest.approx(-2 * math.atan(t), rel=1e-12)  # positive charge rotates clockwise about +z


def test_boris_gyration_radius():
    Bz, gamma = 5e-10, 3.0
    pmag = M_P * C_LIGHT * math.sqrt(gamma ** 2 - 1)
    p = np.array([0.0, pmag, 0.0])
    x = np.zeros(3)
    Omega = Q_E * Bz / (gamma * M_P)
    dt = 2 * math.pi / Omega / 2000
    pts = []
    for _ in range(2000):
        p, x = oracle.boris_push(p, x, (0, 0, 0), (0, 0, Bz), dt)
        pts.append(x.copy())
    pts = np.array(pts)
    r = 0.5 * (pts[:, 0].max() - pts[:, 0].min())
    assert r == pytest.approx(pmag / (Q_E * Bz), rel=1e-5)


def test_boris_ExB_drift():
    B, E = 5e-10, 1e-5  # u x B with u ~ 20 km/s
    gamma = 1.5
    p = np.zeros(3)
    p[0] = M_P * C_LIGHT * math.sqrt(gamma ** 2 - 1)
    x = np.zeros(3)
    Omega = Q_E * B / (gamma * M_P)
    dt = 2 * math.pi / Omega / 200
    nper = 40
    for _ in range(200 * nper):
        p, x = oracle.boris_push(p, x, (E, 0, 0), (0, 0, B), dt)
    vd = E / B  # E x B / B^2 = (E x) x (B z) / B^2 = -E/B y
    vy = x[1] / (200 * nper * dt)
    assert vy == pytest.approx(-vd, rel=0.05)


def test_wrap():
    assert oracle.wrap(1.25 * L, 0.0, L) == pytest.approx(0.25 * L)
    assert oracle.wrap(-0.25 * L, 0.0, L) == pytest.approx(0.75 * L)
    assert oracle.wrap(0.5 * L, 0.0, L) == 0.5 * L


def test_run_cadence_and_kill():
    """save indices are 1:stride:nsteps, no push at si == nsteps, and an open
    boundary kills exactly when x leaves [0, L]."""
    n = 5
    g = oracle.make_geom(n, n, n, L)
    F = uniform_fields(n, B=(0.0, 0.0, 1e-30))  # effectively field-free: straight lines
    gamma = 2.0
    v = C_LIGHT * math.sqrt(1 - 1 / gamma ** 2)
    x = np.array([[0.5 * L], [0.5 * L], [0.5 * L]])
    p = np.array([[gamma * M_P * v], [0.0], [0.0]])
    dt = 0.1 * L / v
    o = oracle.run(F, g, x, p, dt, 1, 12, 12, 5, periodic=False, out_kind=1)
    assert o["nsaved"] == 3  # si = 1, 6, 11
    assert o["pos"][0, 0, 0] == 0.5 * L
    # x crosses L after 5 pushes (0.5 + 0.1 k > 1 at k = 6): alive counts drop at step 6
    assert list(o["alive_counts"][:5]) == [1] * 5 and list(o["alive_counts"][5:]) == [0] * 7
    assert np.isnan(o["pos"][2, 0, 0])  # save at si = 11 is NaN (dead)


# E = -u x B inside the step loop (phase_space_gpu_runner.jl:365-368): pinned
# through oracle.run with a uniform flow, not through boris_push's E argument.
# In B = B0 z and u = u0 x the motional field is E = -u x B = u0 B0 y, whose
# E x B drift E x B / B^2 = +u0 x carries plasma and particles along with the
# flow.  A flipped sign, or the operands swapped (B x u), drifts at -u0 x.
DRIFT_B0, DRIFT_U0 = 5e-10, 3.0e6


def drift_case(n_gyr=40, steps_per_gyr=100, gamma=1.0001, stride=10):
    n = 5
    Lb = 1.0e12
    g = oracle.make_geom(n, n, n, Lb)
    F = uniform_fields(n, B=(0.0, 0.0, DRIFT_B0), u=(DRIFT_U0, 0.0, 0.0))
    vmag = C_LIGHT * math.sqrt(1 - 1 / gamma ** 2)
    # particle 0: gyrating (p perpendicular to B); particle 1: at rest in the
    # flow frame, v = u0 x, where E + v x B = 0 and the Boris step applies no force
    g1 = 1 / math.sqrt(1 - (DRIFT_U0 / C_LIGHT) ** 2)
    x = np.array([[0.4 * Lb, 0.4 * Lb], [0.5 * Lb, 0.5 * Lb], [0.5 * Lb, 0.5 * Lb]])
    p = np.array([[0.0, g1 * M_P * DRIFT_U0], [gamma * M_P * vmag, 0.0], [0.0, 0.0]])
    T_g = 2 * math.pi * gamma * M_P / (Q_E * DRIFT_B0)
    dt = T_g / steps_per_gyr
    nsteps = n_gyr * steps_per_gyr + 1
    return F, g, x, p, dt, nsteps, stride


def test_run_motional_field_drifts_with_the_flow():
    """Uniform B = B0 z, uniform u = u0 x through oracle.run: the gyrating
    particle's guiding centre (least-squares line through every cached
    position over 40 gyrations) moves at +u0 x within 2 %, with no y drift;
    the particle that starts at v = u0 x keeps that velocity (no net force)."""
    F, g, x, p, dt, nsteps, stride = drift_case()
    o = oracle.run(F, g, x, p, dt, 1, nsteps, nsteps, stride, periodic=False, out_kind=1)
    t = np.arange(o["nsaved"]) * stride * dt
    pos = o["pos"]  # (nsave, 3, n)
    vx = np.polyfit(t, pos[:, 0, 0], 1)[0]
    vy = np.polyfit(t, pos[:, 1, 0], 1)[0]
    assert vx == pytest.approx(DRIFT_U0, rel=0.02)
    assert abs(vy) <= 0.02 * DRIFT_U0
    # at rest in the flow frame: a straight line at u0 with momentum kept.
    # The relativistic Boris step balances E + v x B only to O((u0/c)^2 dt)
    # per step (Vay 2008), ~5e-8 of the path here; a flipped E would instead
    # gyrate at 2 u0 with radius 2 u0 / Omega = 1.6 % of the path.
    path = DRIFT_U0 * t[-1]
    assert np.max(np.abs(pos[:, 0, 1] - (0.4 * g.xmax + DRIFT_U0 * t))) <= 1e-6 * path
    assert np.max(np.abs(pos[:, 1, 1] - 0.5 * g.xmax)) <= 1e-6 * path
    assert np.max(np.abs(o["mom"][:, 0, 1] / p[0, 1] - 1.0)) <= 1e-6


def test_run_drift_check_fails_for_a_flipped_motional_field():
    """The same check fed E = +u x B (by reversing u) drifts at -u0 x: the
    drift test above separates the two signs."""
    F, g, x, p, dt, nsteps, stride = drift_case()
    F[3] = -F[3]
    o = oracle.run(F, g, x, p, dt, 1, nsteps, nsteps, stride, periodic=False, out_kind=1)
    t = np.arange(o["nsaved"]) * stride * dt
    vx = np.polyfit(t, o["pos"][:, 0, 0], 1)[0]
    assert vx == pytest.approx(-DRIFT_U0, rel=0.02)


def test_mu_cache_is_mu_of_the_pre_push_phase_cache():
    """out_kind=2 (reconstruct_mu_step_kernel! at save steps,
    multimode_multienergy_cache_pipeline.jl:1349-1374) equals
    reconstruct_mu_chunk (compute_delta_mu2_curve.jl:112-149) applied to the
    same run's out_kind=1 cache, which holds the state before each save
    step's push (test_run_cadence_and_kill pins that); the mu of the
    post-push state differs."""
    from particle_transport_amd import fields
    n_nodes, n = 9, 64
    Lb = fields.box_length_m(10.0)
    raw = fields.generate_turbulence(n_nodes, seed=4, kmax=3, dB_over_B0=1.0)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    g = oracle.make_geom(n_nodes, n_nodes, n_nodes, Lb)
    rng = np.random.default_rng(9)
    x = rng.random((3, n)) * Lb
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    p = d * 1e3 * M_P * C_LIGHT  # gamma ~ 1000: about a cell per 10 pushes at this dt
    dt = 2 * math.pi * 1e3 * M_P / (Q_E * 5e-10) / 20
    total, stride = 201, 10
    o1 = oracle.run(F, g, x.copy(), p.copy(), dt, 1, total, total, stride, True, 1)
    o2 = oracle.run(F, g, x.copy(), p.copy(), dt, 1, total, total, stride, True, 2)
    mu_phase = oracle.reconstruct_mu_chunk(o1["pos"], o1["mom"], F[:3], g)  # (n, nsave)
    assert o2["mu"].shape == (o1["nsaved"], n)
    assert np.array_equal(o2["mu"], mu_phase.T)
    # the state after one more push gives other values
    o3 = oracle.run(F, g, x.copy(), p.copy(), dt, 1, total, total, stride, True, 1)
    xa, pa = x.copy(), p.copy()
    oracle.run(F, g, xa, pa, dt, 1, 1, total, 1, True, 0)
    post = oracle.reconstruct_mu_chunk(xa[None], pa[None], F[:3], g)[:, 0]
    assert not np.allclose(post, o2["mu"][0], rtol=0, atol=1e-9)
    assert np.array_equal(o3["pos"], o1["pos"])


# ----------------------------------------------------------------------- mu

def test_mu_closed_forms():
    assert oracle.mu_from_pB((1, 0, 0), (2, 0, 0)) == 1.0
    assert oracle.mu_from_pB((1, 0, 0), (-2, 0, 0)) == -1.0
    assert oracle.mu_from_pB((0, 1, 0), (3, 0, 0)) == 0.0
    assert oracle.mu_from_pB((1, 1, 0), (1, 0, 0)) == pytest.approx(1 / math.sqrt(2), rel=1e-15)
    assert math.isnan(oracle.mu_from_pB((0, 0, 0), (1, 0, 0)))
    assert math.isnan(oracle.mu_from_pB((float("nan"), 0, 0), (1, 0, 0)))


# -------------------------------------------------------------- statistics

def brute_dmumu(mu, lags, edges, injection, bin_abs):
    """Plain-Python restatement over a tiny (chunk, nsave) series."""
    nb = len(edges) - 1
    res = dict(cnt=np.zeros((len(lags), nb), np.int64), s1=np.zeros((len(lags), nb)), s2=np.zeros((len(lags), nb)),
               pvals=[[] for _ in lags], pss=np.zeros(len(lags)), pc=np.zeros(len(lags), np.int64))
    for li, lag in enumerate(lags):
        if lag >= mu.shape[1]:
            continue
        for row in mu:
            starts = [0] if injection else range(mu.shape[1] - lag)
            vals = []
            for s in starts:
                a, b = float(row[s]), float(row[s + lag])
                if not (math.isfinite(a) and math.isfinite(b)):
                    continue
                d = b - a
                vals.append(d * d)
                v = abs(a) if bin_abs else a
                bi = host_dmumu.mu_bin_index_full(v, edges)
                if bi:
                    res["cnt"][li, bi - 1] += 1
                    res["s1"][li, bi - 1] += d
                    res["s2"][li, bi - 1] += d * d
            if vals:
                res["pvals"][li].append(sum(vals) / len(vals))
                res["pss"][li] += sum(vals)
                res["pc"][li] += len(vals)
    return res


@pytest.mark.parametrize("injection", [False, True])
@pytest.mark.parametrize("bin_abs", [False, True])
def test_dmumu_oracle_vs_brute_force(injection, bin_abs):
    rng = np.random.default_rng(7)
    mu = rng.uniform(-1, 1, (9, 14)).astype(np.float32)
    mu[2, 5] = np.nan
    mu[4, 0] = 1.0
    mu[5, 3] = 0.5
    lags = [1, 3, 7, 13, 20]
    edges = np.linspace(0.0, 1.0, 5) if bin_abs else np.linspace(-1.0, 1.0, 9)
    acc = oracle.new_dmumu_accumulators(len(lags), len(edges) - 1)
    oracle.dmumu(mu, lags, edges, injection, bin_abs, acc)
    b = brute_dmumu(mu, lags, edges, injection, bin_abs)
    assert np.array_equal(acc["dmumu_counts"], b["cnt"])
    assert np.allclose(acc["dmumu_sum_delta"], b["s1"], rtol=1e-14, atol=1e-15)
    assert np.allclose(acc["dmumu_sum_delta2"], b["s2"], rtol=1e-14, atol=1e-15)
    assert np.array_equal(acc["pair_counts"], b["pc"])
    assert np.allclose(acc["pair_sum_squares"], b["pss"], rtol=1e-14)
    for li, vals in enumerate(b["pvals"]):
        if not vals:
            assert acc["particle_counts"][li] == 0
            continue
        v = np.array(vals)
        assert acc["particle_counts"][li] == len(v)
        assert acc["particle_means"][li] == pytest.approx(v.mean(), rel=1e-13)
        # Welford M2 against the two-pass sum of squared deviations
        assert acc["particle_m2s"][li] == pytest.approx(((v - v.mean()) ** 2).sum(), rel=1e-10, abs=1e-15)


def test_dmumu_injection_equals_sliding_at_two_samples():
    mu = np.random.default_rng(3).uniform(-1, 1, (30, 2)).astype(np.float32)
    edges = np.linspace(-1, 1, 11)
    a = oracle.new_dmumu_accumulators(1, 10)
    b = oracle.new_dmumu_accumulators(1, 10)
    oracle.dmumu(mu, [1], edges, True, False, a)
    oracle.dmumu(mu, [1], edges, False, False, b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_welford_chunk_merge_equals_single_pass():
    mu = np.random.default_rng(8).uniform(-1, 1, (40, 30)).astype(np.float32)
    edges = np.linspace(-1, 1, 5)
    one = oracle.new_dmumu_accumulators(2, 4)
    oracle.dmumu(mu, [2, 5], edges, False, False, one)
    two = oracle.new_dmumu_accumulators(2, 4)
    oracle.dmumu(mu[:17], [2, 5], edges, False, False, two)
    oracle.dmumu(mu[17:], [2, 5], edges, False, False, two)
    assert np.array_equal(one["particle_counts"], two["particle_counts"])
    assert np.allclose(one["particle_means"], two["particle_means"], rtol=1e-13)
    assert np.allclose(one["particle_m2s"], two["particle_m2s"], rtol=1e-12)


def test_dpp_oracle_vs_brute_force():
    rng = np.random.default_rng(5)
    mom = rng.normal(0, 1e-18, (11, 3, 6)).astype(np.float32) + np.float32(5e-17)
    mom[4, 1, 2] = np.nan
    lags = [1, 4, 10, 11]
    p0 = 5.2e-17
    acc = oracle.new_dpp_accumulators(len(lags))
    oracle.dpp(mom, lags, p0, acc)
    for li, lag in enumerate(lags):
        c, s1, s2, s3, s4 = 0, 0.0, 0.0, 0.0, 0.0
        if lag < 11:
            for p in range(6):
                for s in range(11 - lag):
                    a = math.sqrt(sum(float(mom[s, k, p]) ** 2 for k in range(3)))
                    b = math.sqrt(sum(float(mom[s + lag, k, p]) ** 2 for k in range(3)))
                    if not (math.isfinite(a) and math.isfinite(b)):
                        continue
                    d = b - a
                    c += 1
                    s1 += d
                    s2 += d * d
                    s3 += d / p0
                    s4 += (d / p0) ** 2
        assert acc["counts"][li] == c
        assert acc["sum_delta_p"][li] == pytest.approx(s1, rel=1e-12, abs=1e-40)
        assert acc["sum_delta_p2"][li] == pytest.approx(s2, rel=1e-12, abs=1e-60)
        assert acc["sum_delta_p_norm2"][li] == pytest.approx(s4, rel=1e-12, abs=1e-30)


def test_kinetic_energy_closed_form():
    gamma = 11.0
    pm = M_P * C_LIGHT * math.sqrt(gamma ** 2 - 1)
    e = (gamma - 1) * M_P * C_LIGHT ** 2 / (1e9 * 1.602176634e-19)
    assert oracle.kinetic_energy_gev(pm) == pytest.approx(e, rel=1e-14)
    assert math.isnan(oracle.kinetic_energy_gev(float("nan")))


def test_mean_abs_b_known_answer():
    # runtests.jl:221-225
    b = [np.full((2, 2, 2), v) for v in (3.0, 4.0, 0.0)]
    assert oracle.mean_abs_b(*b) == 5.0


def test_bin_index_matches_reference_rules():
    edges = np.linspace(0.0, 1.0, 5)
    assert oracle.mu_bin_index(0.0, edges) == 1
    assert oracle.mu_bin_index(0.25, edges) == 2
    assert oracle.mu_bin_index(1.0, edges) == 4  # right-closed last bin
    assert oracle.mu_bin_index(1.0000001, edges) == 0
    assert oracle.mu_bin_index(-1e-9, edges) == 0
    assert oracle.mu_bin_index(float("nan"), edges) == 0
    for v in np.random.default_rng(2).uniform(-0.2, 1.2, 200):
        assert oracle.mu_bin_index(v, edges) == host_dmumu.mu_bin_index_full(v, edges)
