"""The two batched configurations at their full size, through the serving
loop (pt_serve_*).

Mode-decomposed batch: four co-resident 256^3-cell FP32 fields (the total
field and its Alfven/fast/slow parts, dB/B0 = 0.5, k = [1, 16], seed 555),
the clock from the total field, 6 energies log-spaced 1 GeV..100 TeV, fixed
mu0 in {0.1, 0.5, 0.9} and isotropic injection, uneven particle counts 2^14 to
2^20: 96 campaigns submitted together and packed into the same ticks.
Checks: every campaign finishes, the budget is never exceeded, each product
has N (nsave - lag) pairs per lag, and three campaigns (different field,
energy and injection) give bit-identical caches alone and in the batch.

Serving stream: four resident 512^3-cell BF16 fields (dB/B0 = 1, k = [1, 64],
seeds 3000..3003), the first 200 requests of the seeded Poisson stream
(20/s; energy log-uniform 1 GeV..100 TeV, n log-uniform 1e4..1e6, duration
10..500 T_g0, isotropic or fixed-mu injection, a mu cache with D_mumu over
64 log lags x 20 |mu| bins or a D_pp cache), 200 GB budget.  Checks: no
failed request, the reserved budget never above 200 GB, and every product's
pair counts equal N (nsave - lag) for each lag (periodic box: no particle
dies, every mu and |p| is finite)."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

import gpu_fields
from particle_transport_amd import fields, serving
from particle_transport_amd.injection import M_P, Q_E, energy_to_gamma
from particle_transport_amd.runner import Context, Field

pytestmark = pytest.mark.gpu

L_BOX = fields.box_length_m(10.0)
BUDGET = 200 * 10 ** 9


def expected_pairs(nsave, n, lags):
    return np.array([n * (nsave - lag) for lag in lags], dtype=np.int64)


def check_product(res, req, nsave):
    lags = serving.log_lag_steps(nsave, 64)
    want = expected_pairs(nsave, req.n, lags)
    if req.out_kind == serving.OUT_MU:
        got = np.asarray(res.dmumu_counts).sum(axis=1)  # every |mu| falls in [0, 1]
        assert np.array_equal(np.asarray(res.pair_counts), want), req.i
    else:
        got = np.asarray(res.counts)
    assert np.array_equal(got, want), req.i


def mode_requests(B0, mode_gp=10):
    rng = np.random.default_rng(96)
    reqs = []
    for fid in range(4):
        for E in np.geomspace(1.0, 1e5, 6):
            for mu0 in (0.1, 0.5, 0.9, None):
                gamma = float(energy_to_gamma(E))
                dt = (2 * math.pi / 200) / (Q_E * B0 / (gamma * M_P))
                n = int(2 ** rng.uniform(14, 20))
                reqs.append(SimpleNamespace(i=len(reqs), arrival=0.0, field_id=fid, energy_gev=float(E), n=n,
                                            mu0=mu0, out_kind=serving.OUT_MU, dt=dt,
                                            total_steps=mode_gp * 200 + 1, stride=20, gamma=gamma,
                                            seed=int(rng.integers(1 << 31))))
    return reqs


@pytest.mark.timeout(900)
def test_mode_decomposed_batch_96_campaigns():
    gen = gpu_fields.generate(257, seed=555, dB=0.5, kmax=16, modes=True)
    ctx, sctx = Context(0), Context(0)
    flds = [Field(ctx, F, L_BOX, "f32_brick") for F in gen]
    del gen
    B0 = flds[0].B0  # the total field's clock is shared by every mode
    reqs = mode_requests(B0)
    assert len(reqs) == 96 and min(r.n for r in reqs) >= 2 ** 14 and max(r.n for r in reqs) <= 2 ** 20
    probe = {reqs[5].i, reqs[30].i, reqs[71].i}  # fields 0, 1, 2; energies 10 GeV, 1 TeV, 1 TeV; mixed injection
    batch = serving.run_stream(ctx, sctx, flds, reqs, budget_bytes=BUDGET, box_length_m=L_BOX,
                               keep_results=probe, keep_products=True)
    assert batch.failed == [] and batch.requests == 96
    assert batch.peak_reserved_bytes <= BUDGET
    nsave = (reqs[0].total_steps - 1) // reqs[0].stride + 1
    for r in reqs:
        check_product(batch.products[r.i], r, nsave)
    for i in sorted(probe):
        alone = serving.run_stream(ctx, sctx, flds, [reqs[i]], budget_bytes=BUDGET, box_length_m=L_BOX,
                                   keep_results=True)
        a, b = alone.results[i][1], batch.results[i][1]
        assert a.shape == b.shape
        assert np.array_equal(a, b, equal_nan=True), f"campaign {i} differs alone vs in the batch"
    for f in flds:
        f.close()
    sctx.close()
    ctx.close()


@pytest.mark.timeout(900)
def test_serving_stream_first_200_requests():
    ctx, sctx = Context(0), Context(0)
    flds = []
    for k in range(4):
        F = gpu_fields.generate(513, seed=3000 + k, dB=1.0, kmax=64)[0]
        flds.append(Field(ctx, F, L_BOX, "bf16_brick"))
        del F
    reqs = serving.make_stream(200, 4, [f.B0 for f in flds], rate_hz=20.0, seed=77)
    res = serving.run_stream(ctx, sctx, flds, reqs, budget_bytes=BUDGET, box_length_m=L_BOX, keep_products=True)
    print(f"\n200 requests: p50 {res.latency_p50_s:.2f} s, p99 {res.latency_p99_s:.2f} s, "
          f"{res.particle_steps_per_s:.3g} particle-steps/s, peak budget {res.peak_reserved_bytes / 1e9:.1f} GB")
    assert res.failed == [] and res.requests == 200
    assert res.peak_reserved_bytes <= BUDGET
    for r in reqs:
        nsave = (r.total_steps - 1) // r.stride + 1
        check_product(res.products[r.i], r, nsave)
    for f in flds:
        f.close()
    sctx.close()
    ctx.close()
