"""Device injection (pt_inject, pt_ens_create_injected) against the oracle's
init_particles restatement with the same Philox4x64-10 draws
(phase_space_gpu_runner.jl:196-245).  Bars: positions bit-exact (FP64, no
fused multiply-adds on either side); momenta within 1e-15 * |p| (device and
host cos/sin may differ in the last bit); fixed-mu pitch within 1e-13 of mu0
on an FP64 field."""
import ctypes

import numpy as np
import pytest

import oracle
from particle_transport_amd import _lib, fields, injection, runner
from particle_transport_amd.runner import Ensemble, Field

pytestmark = pytest.mark.gpu

L_BOX = fields.box_length_m(10.0)


def turbulent(n_nodes=17, seed=4):
    raw = fields.generate_turbulence(n_nodes, seed=seed, kmax=4, dB_over_B0=1.0)["total"]
    return [np.asfortranarray(a) for a in fields.to_si(raw)]


def params(seed, *, fixed_position=False, position=(0.0, 0.0, 0.0), fixed_mu=False, mu0=0.0, pscale=1.0):
    ip = _lib.InjectParams()
    ip.seed = seed
    ip.fixed_position = int(fixed_position)
    ip.fixed_mu = int(fixed_mu)
    for c in range(3):
        ip.position_m[c] = position[c]
    ip.mu0 = mu0
    ip.p_magnitude = pscale
    return ip


def device_inject(ctx, fld, ip, first, n, on_device=False):
    L = _lib.load()
    if not on_device:
        out = np.empty((6, n))
        _lib.check(L.pt_inject(ctx.h, fld.h, first, n, ctypes.byref(ip), 0,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out
    buf = _lib.DeviceBuffer(ctx, (6, n), np.float64)
    _lib.check(L.pt_inject(ctx.h, fld.h, first, n, ctypes.byref(ip), 1,
                           ctypes.cast(ctypes.c_void_p(buf.ptr), ctypes.POINTER(ctypes.c_double))))
    out = buf.to_host()
    buf.free()
    return out


MODES = [dict(), dict(fixed_mu=True, mu0=0.37), dict(fixed_position=True, position=(0.3 * L_BOX, L_BOX, 0.0)),
         dict(fixed_position=True, position=(0.61 * L_BOX, 0.2 * L_BOX, 0.9 * L_BOX), fixed_mu=True, mu0=-0.8)]


@pytest.mark.parametrize("mode", range(len(MODES)))
def test_inject_matches_oracle(ctx, mode):
    F = turbulent()
    g = oracle.make_geom(17, 17, 17, L_BOX)
    fld = Field(ctx, F, L_BOX, "f64_planar")
    kw = dict(MODES[mode], pscale=3.2e-16)
    n, first, seed = 4099, 123, 2024
    got = device_inject(ctx, fld, params(seed, **kw), first, n)
    ref, st = oracle.inject(seed, first, n, g, B=F[:3], **kw)
    assert st == 0
    assert np.array_equal(got[:3], ref[:3])  # positions: bit-exact
    assert np.max(np.abs(got[3:] - ref[3:])) <= 1e-15 * kw["pscale"]
    if kw.get("fixed_mu"):
        for q in range(0, n, 97):
            b = [oracle.trilinear(F[c], *got[:3, q], g) for c in range(3)]
            assert abs(oracle.mu_from_pB(got[3:, q], b) - kw["mu0"]) < 1e-13
    fld.close()


def test_inject_split_and_device_output_are_identical(ctx):
    F = turbulent()
    fld = Field(ctx, F, L_BOX, "f32_brick")
    ip = params(7, fixed_mu=True, mu0=0.5, pscale=1.0)
    whole = device_inject(ctx, fld, ip, 0, 10_000)
    parts = np.concatenate([device_inject(ctx, fld, ip, a, b - a) for a, b in ((0, 4097), (4097, 10_000))], axis=1)
    assert np.array_equal(whole, parts)
    assert np.array_equal(whole, device_inject(ctx, fld, ip, 0, 10_000, on_device=True))
    fld.close()


def test_ensemble_created_on_device_holds_the_injected_state(ctx):
    F = turbulent()
    fld = Field(ctx, F, L_BOX, "f64_planar")
    ip = params(31, pscale=5.0e-17)
    n = 3001
    ens = Ensemble.injected(ctx, fld, n, ip, pusher="ref_f64", periodic=True)
    x, p, alive = ens.state()
    ref = device_inject(ctx, fld, ip, 0, n)
    assert np.array_equal(x, ref[:3]) and np.array_equal(p, ref[3:]) and alive.all()
    # the same state through the host path gives the same ensemble
    ens2 = Ensemble(ctx, fld, ref[:3], ref[3:], pusher="ref_f64")
    x2, p2, _ = ens2.state()
    assert np.array_equal(x, x2) and np.array_equal(p, p2)
    ens.close()
    ens2.close()
    fld.close()


@pytest.mark.parametrize("storage", ["f16_brick", "f32_brick"])
def test_fixed_mu_on_device_gives_mu0_in_the_pushers_field(ctx, storage):
    """The first cadence sample (si = 1, before any push) of the production
    pusher's mu cache is mu0 about the field the pusher interpolates."""
    F = turbulent()
    fld = Field(ctx, F, L_BOX, storage)
    n, mu0 = 8192, 0.42
    ens = Ensemble.injected(ctx, fld, n, params(5, fixed_mu=True, mu0=mu0, pscale=1.0e-17), pusher="fast")
    r = ens.advance(dt=1e-3, first_step=1, n_steps=1, total_steps=10, stride=1, out_kind="mu")
    mu = r.out_a[0]
    assert np.all(np.isfinite(mu))
    assert np.max(np.abs(mu - mu0)) < 2e-6
    ens.close()
    fld.close()


def test_fixed_mu_zero_field_is_an_error(ctx):
    F = [np.zeros((9, 9, 9), order="F") for _ in range(6)]
    fld = Field(ctx, F, L_BOX, "f64_planar")
    # isotropic injection needs no field values (this also sizes the context scratch)
    device_inject(ctx, fld, params(1), 0, 100)
    ens = Ensemble.injected(ctx, fld, 100, params(1), pusher="ref_f64")
    ens.close()
    before = ctx.bytes_in_use()
    with pytest.raises(_lib.PTError, match="local B is zero"):
        Ensemble.injected(ctx, fld, 100, params(1, fixed_mu=True, mu0=0.5), pusher="ref_f64")
    with pytest.raises(_lib.PTError, match="local B is zero"):
        device_inject(ctx, fld, params(1, fixed_mu=True, mu0=0.5), 0, 100)
    assert ctx.bytes_in_use() == before  # the failed ensemble released its buffers
    with pytest.raises(_lib.PTError, match="outside the simulation box"):
        device_inject(ctx, fld, params(1, fixed_position=True, position=(0.0, 1.01 * L_BOX, 0.0)), 0, 10)
    with pytest.raises(_lib.PTError, match="injection_mu0"):
        device_inject(ctx, fld, params(1, fixed_mu=True, mu0=1.5), 0, 10)
    fld.close()


def test_large_isotropic_injection_moments(ctx):
    """4,194,304 particles on the device: positions inside the box, isotropic
    moments within 5 sigma."""
    F = turbulent()
    fld = Field(ctx, F, L_BOX, "f16_brick")
    n = 4_194_304
    out = device_inject(ctx, fld, params(99, pscale=1.0), 0, n, on_device=True)
    assert np.all((out[:3] >= 0) & (out[:3] <= L_BOX))
    uz = out[5]
    assert abs(uz.mean()) < 5 * np.sqrt(1 / 3 / n)
    assert abs((uz ** 2).mean() - 1 / 3) < 5 * np.sqrt(4 / 45 / n)
    assert np.allclose(np.sqrt((out[3:] ** 2).sum(axis=0)), 1.0, rtol=2e-15, atol=0)
    ref, _ = oracle.inject(99, n - 1000, 1000, oracle.make_geom(17, 17, 17, L_BOX), pscale=1.0)
    assert np.array_equal(out[:3, n - 1000:], ref[:3])
    fld.close()


def test_runner_with_device_injection():
    """run_mu_cache_ensemble with injection_rng = "device" equals the same
    run handed the pt_inject state as its initial state, and records the
    injection settings (write_injection_metadata!, :136-143)."""
    F = turbulent()
    cfg = dict(box_length_pc=10.0, boundary="periodic", precision="Float64", integration_steps_per_gyroperiod=50.0,
               trajectory_duration_gyroperiods=2.01, trajectory_save_interval_gyroperiods=0.2, use_cfl=False,
               cfl=0.2, eta=0.1, tOmega0_max=1.0, trajectory_time_stride=1, seed=17, n_particles=500,
               trajectory_mode="total", injection_mode="fixed-mu", injection_mu0=0.6,
               mu_cache_output_precision="Float64", progress_every=0, injection_rng="device")
    with runner.Context(0) as c:
        tr = runner.resolve_total_field_time_reference(cfg, F, energy_GeV=10.0, ctx=c)
        fld = Field(c, F, L_BOX, "f64_planar")
        a = runner.run_mu_cache_ensemble(cfg, F, energy_GeV=10.0, time_reference=tr, ctx=c, field=fld)
        s = runner._trajectory_setup(cfg, F, 10.0, tr, "test")
        state = injection.init_particles_device(cfg, c, fld, s.axes, s.gamma0, s.v0)
        b = runner.run_mu_cache_ensemble(cfg, F, energy_GeV=10.0, time_reference=tr, ctx=c, field=fld,
                                         initial_state=state)
        fld.close()
    mu_a, mu_b = np.asarray(a.cache["mu"]), np.asarray(b.cache["mu"])
    assert np.array_equal(mu_a, mu_b)
    assert np.max(np.abs(mu_a[0] - 0.6)) < 1e-13  # first cadence sample is the injected pitch
    assert a.cache.meta["injection_mode"] == "fixed_mu" and a.cache.meta["injection_rng"] == "device"
    assert a.cache.meta["injection_mu0"] == [0.6]
