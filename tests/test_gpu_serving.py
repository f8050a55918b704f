"""GPU tests of the continuous-batching server: a request's results are the
same alone and in a mix (mu cache bit-identical, D_mumu sums within 1e-12),
equal the single-campaign production pusher, admission respects the HBM
budget, and fixed-mu injection starts at mu0."""
import math

import numpy as np
import pytest

from particle_transport_amd import fields, serving
from particle_transport_amd.injection import C_LIGHT, M_P, Q_E, energy_to_gamma
from particle_transport_amd.runner import Context, Ensemble, Field

pytestmark = pytest.mark.gpu

L_BOX = fields.box_length_m(10.0)


@pytest.fixture(scope="module")
def setup():
    ctx = Context(0)
    stats = Context(0)
    raw = fields.generate_turbulence(33, seed=21, kmax=8, dB_over_B0=1.0)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    fld = Field(ctx, F, L_BOX, "bf16_brick")
    yield ctx, stats, fld, F
    fld.close()
    stats.close()
    ctx.close()


def req(i, n, E, dur_steps, kind, B0, mu0=None, seed=0):
    gamma = float(energy_to_gamma(E))
    dt = (2 * math.pi / 100) / (Q_E * B0 / (gamma * M_P))
    from types import SimpleNamespace
    return SimpleNamespace(i=i, arrival=0.0, field_id=0, energy_gev=E, n=n, mu0=mu0, out_kind=kind, dt=dt,
                           total_steps=dur_steps, stride=10, gamma=gamma, seed=seed)


def run_alone(ctx, fld, r, steps_per_tick=64, on_device=True):
    srv = serving.Server(ctx, 1 << 32, steps_per_tick)
    srv.add_field(fld)
    rid = serving.submit_request(srv, r, L_BOX, on_device=on_device)
    while srv.poll(rid) != serving.DONE:
        srv.step()
    out = srv.copy_cache(rid, serving.cache_dtype(r.out_kind))
    _, _, _, alive = srv.fetch(rid)
    srv.release(rid)
    srv.close()
    return out, alive


@pytest.mark.parametrize("on_device", [True, False], ids=["device_injection", "host_state"])
def test_alone_equals_mix_and_single_campaign_pusher(setup, on_device):
    ctx, stats, fld, F = setup
    B0 = fld.B0
    A = req(0, 3000, 10.0, 401, serving.OUT_MU, B0, seed=1)
    B = req(1, 777, 1e4, 251, serving.OUT_PMAG, B0, seed=2)
    C = req(2, 5000, 3.0, 333, serving.OUT_MU, B0, mu0=0.4, seed=3)
    alone = {r.i: run_alone(ctx, fld, r, on_device=on_device) for r in (A, B, C)}
    srv = serving.Server(ctx, 1 << 32, 64)
    srv.add_field(fld)
    ids = {serving.submit_request(srv, A, L_BOX, on_device=on_device): A}
    srv.step()
    srv.step()  # B and C join while A is running
    ids[serving.submit_request(srv, B, L_BOX, on_device=on_device)] = B
    srv.step()
    ids[serving.submit_request(srv, C, L_BOX, on_device=on_device)] = C
    mixed = {}
    while len(mixed) < 3:
        srv.step()
        for rid, r in ids.items():
            if r.i not in mixed and srv.poll(rid) == serving.DONE:
                mixed[r.i] = (srv.copy_cache(rid, serving.cache_dtype(r.out_kind)),
                              srv.fetch(rid)[3])
                res_mix = serving.finish_request(srv, stats, rid, r)
                if r.i == 0:
                    mix_acc = res_mix
    srv.close()
    for i in range(3):
        a, b = alone[i][0], mixed[i][0]
        assert a.shape == b.shape
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), f"request {i} differs alone vs mixed"
        assert alone[i][1] == mixed[i][1]
    # D_mumu of request A from its cache: alone vs mixed
    lags = serving.log_lag_steps(alone[0][0].shape[0], 64)
    from particle_transport_amd._lib import DeviceBuffer
    buf = DeviceBuffer.from_host(stats, alone[0][0])
    acc_alone = serving.dmumu_on_device(stats, buf.ptr, alone[0][0].shape[0], 3000, lags, np.linspace(0, 1, 21))
    buf.free()
    # the bin sums are exact fixed-point sums and the Welford merges run in a
    # fixed order: every accumulator is bitwise the same alone and in the mix
    for f in ("dmumu_counts", "dmumu_sum_delta", "dmumu_sum_delta2", "particle_counts", "particle_means",
              "particle_m2s", "pair_sum_squares", "pair_counts"):
        assert np.array_equal(getattr(acc_alone, f), getattr(mix_acc, f)), f"D_mumu {f} differs alone vs mixed"
    # the same requests through the single-campaign production pusher (the
    # device-injected ones from pt_ens_create_injected with the same draws)
    for R in (A, C):
        if on_device:
            ens = Ensemble.injected(ctx, fld, R.n, serving.request_inject_params(R), pusher="fast")
        elif R.mu0 is None:
            x, p, _, _ = serving.request_state(R, L_BOX)
            ens = Ensemble(ctx, fld, x, p, pusher="fast")
        else:
            continue  # host fixed-mu states are built inside the server from host gyrophases
        r = ens.advance(dt=R.dt, first_step=1, n_steps=R.total_steps, total_steps=R.total_steps, stride=R.stride,
                        out_kind="mu", steps_per_launch=100)
        ens.close()
        assert np.array_equal(r.out_a, alone[R.i][0]), R.i


def test_non_cubic_grid_equals_single_campaign_pusher(setup):
    """A field with three cell sizes (33 x 17 x 25 nodes) takes the general
    position step in both the server and the single-campaign pusher: the
    request's mu cache equals the pusher's bit for bit."""
    ctx, stats, _, F = setup
    fld = Field(ctx, [np.asfortranarray(a[:, :17, :25]) for a in F], L_BOX, "bf16_brick")
    R = req(0, 3000, 1e4, 301, serving.OUT_MU, fld.B0, seed=5)
    out, alive = run_alone(ctx, fld, R)
    ens = Ensemble.injected(ctx, fld, R.n, serving.request_inject_params(R), pusher="fast")
    r = ens.advance(dt=R.dt, first_step=1, n_steps=R.total_steps, total_steps=R.total_steps, stride=R.stride,
                    out_kind="mu", steps_per_launch=100)
    ens.close()
    fld.close()
    assert np.isfinite(out).any()
    assert np.array_equal(r.out_a, out, equal_nan=True)


def test_relative_momentum_cache_and_dpp(setup):
    """A D_pp request with the relative cache (PT_OUT_DPREL) holds exactly
    float32(|p|/p_ref - 1) of the same request's FP64 |p| cache, and its D_pp
    agrees with the one from the |p| cache to the float32 rounding of delta."""
    ctx, stats, fld, F = setup
    A = req(0, 4000, 3e4, 301, serving.OUT_PMAG, fld.B0, seed=12)
    B = req(0, 4000, 3e4, 301, serving.OUT_DPREL, fld.B0, seed=12)
    pm, _ = run_alone(ctx, fld, A)
    dr, _ = run_alone(ctx, fld, B)
    p_ref = A.gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / A.gamma ** 2)
    want = (pm / p_ref - 1.0).astype(np.float32)
    assert dr.dtype == np.float32 and dr.shape == pm.shape
    assert np.array_equal(np.isnan(dr), np.isnan(want)) and np.array_equal(dr[~np.isnan(dr)], want[~np.isnan(want)])
    from particle_transport_amd._lib import DeviceBuffer
    lags = serving.log_lag_steps(pm.shape[0], 64)
    b1, b2 = DeviceBuffer.from_host(stats, pm), DeviceBuffer.from_host(stats, dr)
    a1 = serving.dpp_on_device(stats, b1.ptr, pm.shape[0], 4000, lags, p_ref)
    a2 = serving.dpp_dprel_on_device(stats, b2.ptr, pm.shape[0], 4000, lags, p_ref, p_ref)
    b1.free()
    b2.free()
    assert np.array_equal(a1.counts, a2.counts)
    assert np.allclose(a2.sum_delta_p2, a1.sum_delta_p2, rtol=1e-5, atol=0)
    assert np.allclose(a2.sum_delta_p_norm2, a1.sum_delta_p_norm2, rtol=1e-5, atol=0)


def test_budget_admission_fifo(setup):
    ctx, stats, fld, F = setup
    B0 = fld.B0
    A = req(0, 20000, 10.0, 301, serving.OUT_MU, B0, seed=4)
    B = req(1, 20000, 10.0, 301, serving.OUT_MU, B0, seed=5)
    one = 9 * 4 * 20000 + 20000 + 31 * 20000 * 4 + 8192  # state + alive + cache (+ alignment slack)
    srv = serving.Server(ctx, int(1.5 * one), 50)
    srv.add_field(fld)
    a = serving.submit_request(srv, A, L_BOX)
    b = serving.submit_request(srv, B, L_BOX)
    running, queued, _ = srv.step()
    assert (running, queued) == (1, 1) and srv.poll(b) == serving.QUEUED
    while srv.poll(a) != serving.DONE:
        srv.step()
    assert srv.poll(b) == serving.QUEUED  # still waiting: A holds the memory until released
    srv.release(a)
    running, queued, _ = srv.step()
    assert (running, queued) == (1, 0) and srv.poll(b) == serving.RUNNING
    while srv.poll(b) != serving.DONE:
        srv.step()
    srv.release(b)
    assert srv.stats().bytes_reserved == 0
    srv.close()


@pytest.mark.parametrize("on_device", [True, False], ids=["device_injection", "host_state"])
def test_fixed_mu_injection_starts_at_mu0(setup, on_device):
    ctx, stats, fld, F = setup
    r = req(0, 4096, 100.0, 11, serving.OUT_MU, fld.B0, mu0=-0.6, seed=9)
    mu, _ = run_alone(ctx, fld, r, on_device=on_device)
    assert np.max(np.abs(mu[0] + 0.6)) <= 2e-5


def test_device_injection_on_a_zero_field_fails_at_fetch(setup):
    """A fixed-mu request where the local B is zero: admission does not wait
    for the injection kernel, so the error (the reference's "cannot use
    fixed-mu injection where local B is zero") is reported when the request's
    result is fetched; a request beside it is unaffected."""
    ctx, stats, fld, F = setup
    zero = Field(ctx, [np.zeros((9, 9, 9), order="F") for _ in range(6)], L_BOX, "f32_brick")
    srv = serving.Server(ctx, 1 << 30, 50)
    srv.add_field(zero)
    bad = req(0, 1000, 10.0, 101, serving.OUT_MU, 1e-10, mu0=0.3, seed=5)
    good = req(1, 1000, 10.0, 101, serving.OUT_MU, 1e-10, seed=6)
    ib = serving.submit_request(srv, bad, L_BOX)
    ig = serving.submit_request(srv, good, L_BOX)
    while srv.poll(ib) != serving.DONE or srv.poll(ig) != serving.DONE:
        srv.step()
    from particle_transport_amd._lib import PTError
    with pytest.raises(PTError, match="local B is zero"):
        srv.fetch(ib)
    assert srv.fetch(ig)[3] == 1000
    srv.release(ib)
    srv.release(ig)
    srv.close()
    zero.close()


def test_released_requests_return_their_memory(setup):
    """Released requests give their memory back while the server keeps
    running, as in the serving loop: 60 requests arrive together under a
    3 GiB budget (several run at once), each finished request's D_mumu product
    runs on the statistics context in a worker thread beside the next ticks,
    and the request is released when its product is done.  About 25 GB of
    requests pass through the server's arena (the budget plus 12.5 %), so
    device free memory must stay within that arena plus 1 GiB of where it
    started, and every request must have been admitted."""
    from concurrent.futures import ThreadPoolExecutor
    ctx, stats, fld, F = setup
    start, _ = ctx.device_memory()
    srv = serving.Server(ctx, 3 << 30, 200)
    srv.add_field(fld)
    low = start
    rng = np.random.default_rng(8)
    pending = {}
    for i in range(60):
        n = int(rng.integers(40000, 200000))  # cache: 1001 samples x n x 4 B
        r = req(i, n, 10.0, 1001, serving.OUT_MU, fld.B0, seed=200 + i)
        r.stride = 1
        pending[serving.submit_request(srv, r, L_BOX)] = r
    finishing = {}
    with ThreadPoolExecutor(max_workers=1) as pool:
        while pending or finishing:
            srv.step()
            for rid in [k for k in pending if srv.poll(k) == serving.DONE]:
                r = pending.pop(rid)
                cp, nsave, n, _ = srv.fetch(rid)
                finishing[rid] = pool.submit(serving.request_product, stats, cp, nsave, n, r)
            for rid in [k for k, f in finishing.items() if f.done()]:
                finishing.pop(rid).result()
                srv.release(rid)
            low = min(low, ctx.device_memory()[0])
    assert srv.stats().bytes_reserved == 0
    assert start - low < int(1.125 * (3 << 30)) + (1 << 30), (start - low) / 2 ** 30
    srv.close()


def test_oversized_request_rejected(setup):
    ctx, stats, fld, F = setup
    srv = serving.Server(ctx, 1 << 20, 50)
    srv.add_field(fld)
    r = req(0, 100000, 10.0, 1001, serving.OUT_MU, fld.B0)
    from particle_transport_amd._lib import PTError
    with pytest.raises(PTError):
        serving.submit_request(srv, r, L_BOX)
    srv.close()
