"""Device-injection draws and the oracle's init_particles restatement, on the
CPU: the Philox4x64-10 blocks against numpy.random.Philox (an independent
implementation), positions against numpy's doubles mapped onto the box, the
fixed-mu and isotropic closed forms, and the host checks of the settings
(init_particles, phase_space_gpu_runner.jl:196-245)."""
import math

import numpy as np
import pytest

import oracle
from particle_transport_amd import fields, injection
from particle_transport_amd.time_units import PipelineError

L_BOX = fields.box_length_m(10.0)
MASK64 = (1 << 64) - 1


def numpy_block(key, counter):
    """Block `counter` (four words, little-endian) of numpy's Philox: numpy
    increments the counter before each block, so start one below it."""
    c = sum(int(w) << (64 * k) for k, w in enumerate(counter))
    k = int(key[0]) + (int(key[1]) << 64)
    return np.random.Philox(key=k, counter=(c - 1) % (1 << 256))


def test_philox_blocks_match_numpy():
    rng = np.random.default_rng(3)
    cases = [([0, 0, 0, 0], [0, 0]), ([MASK64] * 4, [MASK64] * 2), ([1, 0, 0, 0], [12345, 0])]
    for _ in range(40):
        cases.append(([int(v) for v in rng.integers(0, 1 << 63, 4, dtype=np.uint64)],
                      [int(v) for v in rng.integers(0, 1 << 63, 2, dtype=np.uint64)]))
    for ctr, key in cases:
        expect = [int(v) for v in numpy_block(key, ctr).random_raw(4)]
        assert oracle.philox4x64_10(ctr, key) == expect, (ctr, key)


def test_draws_are_numpy_philox_doubles():
    # draw j of particle i = word j % 4 of block (i, j // 4), as numpy's random() maps it
    for seed in (0, 2024, MASK64):
        for i in (0, 1, 77, (1 << 31) - 1):
            first = np.random.Generator(numpy_block([seed, 0], [i, 0, 0, 0])).random(4)
            second = np.random.Generator(numpy_block([seed, 0], [i, 1, 0, 0])).random(4)
            got = [oracle.inject_draw(seed, i, j) for j in range(8)]
            assert got == list(first) + list(second)


def test_random_positions_and_isotropic_directions_follow_the_draws():
    n, seed, pscale = 300, 99, 3.5e-17
    g = oracle.make_geom(17, 17, 17, L_BOX)
    out, st = oracle.inject(seed, 10, n, g, pscale=pscale)
    assert st == 0
    u = np.array([[oracle.inject_draw(seed, 10 + q, j) for j in range(5)] for q in range(n)])
    for c in range(3):  # rand(rng, T) * (xmax - xmin) + xmin, bit for bit
        assert np.array_equal(out[c], u[:, c] * (L_BOX - 0.0) + 0.0)
    z = 2 * u[:, 3] - 1
    phi = 2 * math.pi * u[:, 4]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    for c, v in enumerate((r * np.cos(phi), r * np.sin(phi), z)):
        assert np.allclose(out[3 + c], pscale * v, rtol=0, atol=4e-16 * pscale)
    assert np.allclose(np.linalg.norm(out[3:], axis=0), pscale, rtol=1e-15)
    assert np.all((out[:3] >= 0) & (out[:3] <= L_BOX))


def test_particle_state_does_not_depend_on_the_split():
    g = oracle.make_geom(9, 9, 9, L_BOX)
    whole, _ = oracle.inject(5, 0, 100, g, pscale=1.0)
    a, _ = oracle.inject(5, 0, 37, g, pscale=1.0)
    b, _ = oracle.inject(5, 37, 63, g, pscale=1.0)
    assert np.array_equal(whole, np.concatenate([a, b], axis=1))
    other, _ = oracle.inject(6, 0, 100, g, pscale=1.0)
    assert not np.any(other[0] == whole[0])


def test_isotropic_moments():
    n = 200_000
    g = oracle.make_geom(9, 9, 9, L_BOX)
    out, _ = oracle.inject(11, 0, n, g, pscale=1.0)
    uz = out[5]
    phi = np.arctan2(out[4], out[3])
    assert abs(uz.mean()) < 5 * math.sqrt(1 / 3 / n)
    assert abs((uz ** 2).mean() - 1 / 3) < 5 * math.sqrt(4 / 45 / n)
    assert abs(np.cos(phi).mean()) < 5 * math.sqrt(0.5 / n)
    assert abs(np.sin(phi).mean()) < 5 * math.sqrt(0.5 / n)
    for c in range(3):  # positions uniform over the box
        assert abs(out[c].mean() / L_BOX - 0.5) < 5 * math.sqrt(1 / 12 / n)


def turbulent_B(n_nodes=17, seed=4):
    raw = fields.generate_turbulence(n_nodes, seed=seed, kmax=4, dB_over_B0=1.0)["total"]
    return [np.asfortranarray(a) for a in fields.to_si(raw)[:3]]


@pytest.mark.parametrize("mu0", [-1.0, -0.3, 0.0, 0.71, 1.0])
def test_fixed_mu_pitch_is_mu0_about_local_B(mu0):
    n, pscale = 500, 2.0e-16
    B = turbulent_B()
    g = oracle.make_geom(17, 17, 17, L_BOX)
    out, st = oracle.inject(8, 0, n, g, fixed_mu=True, mu0=mu0, pscale=pscale, B=B)
    assert st == 0
    for q in range(0, n, 7):
        b = [oracle.trilinear(B[c], *out[:3, q], g) for c in range(3)]
        assert abs(oracle.mu_from_pB(out[3:, q], b) - mu0) < 1e-13
    assert np.allclose(np.linalg.norm(out[3:], axis=0), pscale, rtol=1e-14)
    # fixed mu uses one draw after the positions: the gyrophase 2 pi u
    q = 3
    u3 = oracle.inject_draw(8, q, 3)
    b = np.array([oracle.trilinear(B[c], *out[:3, q], g) for c in range(3)])
    bh = b / np.linalg.norm(b)
    perp = out[3:, q] / pscale - mu0 * bh
    if abs(mu0) < 1:
        e1 = np.array([0.0, bh[2], -bh[1]]) if abs(bh[0]) < 0.9 else np.array([-bh[2], 0.0, bh[0]])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(bh, e1)
        ang = math.atan2(perp @ e2, perp @ e1) % (2 * math.pi)
        assert abs(ang - 2 * math.pi * u3) < 1e-12


def test_fixed_position_and_zero_field():
    g = oracle.make_geom(9, 9, 9, L_BOX)
    pos = (0.25 * L_BOX, L_BOX, 0.0)
    out, st = oracle.inject(1, 0, 20, g, fixed_position=True, position=pos, pscale=1.0)
    assert st == 0 and all(np.all(out[c] == pos[c]) for c in range(3))
    # isotropic with a fixed position uses draws 0 and 1 for the direction
    assert out[5, 4] == 2 * oracle.inject_draw(1, 4, 0) - 1
    zero = [np.zeros((9, 9, 9))] * 3
    _, st = oracle.inject(1, 0, 4, g, fixed_mu=True, mu0=0.5, pscale=1.0, B=zero)
    assert st == -1  # the reference raises "cannot use fixed-mu injection where local B is zero"


def test_inject_params_host_checks():
    axes = [dict(min=0.0, max=L_BOX, d=L_BOX / 8, n=9)] * 3
    cfg = dict(seed=-1, n_particles=10, injection_mode="fixed-mu", injection_mu0=0.4,
               injection_position_mode="fixed", injection_position=(0.5, 0.25, 1.0))
    ip = injection.inject_params(cfg, axes, 2.0, 1.0e8)
    assert ip.seed == MASK64 and ip.fixed_mu == 1 and ip.fixed_position == 1
    assert list(ip.position_m) == [0.5 * L_BOX, 0.25 * L_BOX, L_BOX]
    assert ip.mu0 == 0.4 and ip.p_magnitude == 2.0 * injection.M_P * 1.0e8
    with pytest.raises(PipelineError, match="injection_mu0"):
        injection.inject_params(dict(cfg, injection_mu0=1.5), axes, 2.0, 1.0e8)
    with pytest.raises(PipelineError, match="outside the simulation box"):
        injection.inject_params(dict(cfg, injection_position=(0.5, 0.5, 1.5)), axes, 2.0, 1.0e8)
    with pytest.raises(PipelineError, match="injection_rng"):
        injection.injection_rng(dict(injection_rng="gpu"))
    assert injection.injection_rng({}) == "host" and injection.injection_rng(dict(injection_rng="Device")) == "device"
