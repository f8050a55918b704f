"""The precision-form configuration's accuracy bar, at its field and energy:
a 512^3-cell periodic field (513^3 nodes, dB/B0 = 1.0 over k = 2pi/L [1, 64],
u rms 20 km/s, seed 2024, L = 10 pc), protons at 100 GeV, 200 steps per
T_g0, mu cached every 20 steps over the whole 500 T_g0 campaign (100,000
pushes, 5001 samples), on a 65,536-particle subset.

The production pusher on BF16, FP16 and FP32 bricks runs from the same
initial states and dt as the FP64 reference-structured pusher on the FP64
field (the kernel the parity tests pin to the oracle).  D_mumu(|mu|) uses the
serving products' grid, 64 log-spaced lags over the whole series x 20 |mu|
bins, sliding starts.  Passing means, in every |mu| bin, the count-weighted
tau average of D = <dmu^2>/(2 tau) within max(1 %, 2 sigma) of the FP64 run,
and in every lag the bin-summed <dmu^2>(tau) within max(1 %, 2 sigma); sigma
is the sampling error of the paired difference over 16 particle groups."""
import math

import numpy as np
import pytest

import gpu_fields
from particle_transport_amd import dmumu, fields
from particle_transport_amd._lib import DeviceBuffer
from particle_transport_amd.injection import C_LIGHT, M_P, Q_E, energy_to_gamma
from particle_transport_amd.runner import Context, Ensemble, Field

pytestmark = pytest.mark.gpu

N_SUB, S, STRIDE, T_G0 = 65536, 200, 20, 500
GROUPS = 16


def per_group_curves(ctx, mu_dev, lags, edges):
    """For all particles and each of GROUPS groups: per-bin tau-averaged D and
    per-lag bin-summed <dmu^2>, with the bin pair counts."""
    tau = np.asarray(lags, dtype=np.float64)
    out = []
    for gi in range(GROUPS + 1):
        sel = np.arange(N_SUB) if gi == GROUPS else np.arange(gi * N_SUB // GROUPS, (gi + 1) * N_SUB // GROUPS)
        acc = dmumu.new_accumulators(len(lags), len(edges) - 1)
        dmumu.process_mu_chunk_dmumu(ctx, mu_dev, lags, edges, acc, "sliding", True, layout="particle_fastest",
                                     particle_index=sel)
        w = acc.dmumu_counts  # (lags, bins)
        D = acc.dmumu_sum_delta2 / np.maximum(w, 1) / (2 * tau[:, None])
        d_bin = (D * w).sum(0) / np.maximum(w.sum(0), 1)
        d_lag = acc.dmumu_sum_delta2.sum(1) / np.maximum(w.sum(1), 1)
        out.append((d_bin, d_lag, w.sum(0)))
    return out


def compare(a, b, what):
    k = {"bin": 0, "lag": 1}[what]
    ref = b[GROUPS][k]
    diff = a[GROUPS][k] - ref
    gd = np.array([a[g][k] - b[g][k] for g in range(GROUPS)])
    sigma = gd.std(0, ddof=1) / math.sqrt(GROUPS)
    ok = b[GROUPS][2] >= 1000 if what == "bin" else np.ones(len(ref), bool)
    rel = np.abs(diff[ok]) / np.abs(ref[ok])
    passed = np.abs(diff[ok]) <= np.maximum(0.01 * np.abs(ref[ok]), 2 * sigma[ok])
    return rel, np.abs(diff[ok]) / np.maximum(sigma[ok], 1e-300), passed


@pytest.fixture(scope="module")
def c3_runs():
    F32 = gpu_fields.generate(513, seed=2024, dB=1.0, kmax=64)[0]
    L = fields.box_length_m(10.0)
    rng = np.random.default_rng(99)
    x = rng.random((3, N_SUB)) * L
    d = rng.standard_normal((3, N_SUB))
    d /= np.linalg.norm(d, axis=0)
    gamma = float(energy_to_gamma(100.0))
    p = d * (gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / gamma ** 2))
    steps = T_G0 * S + 1
    nsave = (steps - 1) // STRIDE + 1
    lags = sorted({int(round(v)) for v in np.geomspace(1, nsave - 1, 64)})
    edges = np.linspace(0.0, 1.0, 21)
    curves = {}
    ctx = Context(0)
    F64 = tuple(np.asfortranarray(c, dtype=np.float64) for c in F32)
    fld = Field(ctx, F64, L, "f64_planar")
    B0 = fld.B0
    del F64
    dt = (2 * math.pi / S) / (Q_E * B0 / (gamma * M_P))
    for name, pusher, storage in (("fp64", "ref_f64", None), ("bf16", "fast", "bf16_brick"),
                                  ("f16", "fast", "f16_brick"), ("f32", "fast", "f32_brick")):
        if storage is not None:
            fld = Field(ctx, F32, L, storage)
        ens = Ensemble(ctx, fld, x, p, pusher=pusher)
        mu = DeviceBuffer(ctx, (nsave, N_SUB), np.float32)
        ens.advance(dt=dt, first_step=1, n_steps=steps, total_steps=steps, stride=STRIDE, out_kind="mu", out_a=mu,
                    steps_per_launch=S)
        ens.close()
        fld.close()
        curves[name] = per_group_curves(ctx, mu, lags, edges)
        mu.free()
    ctx.close()
    return curves, lags


@pytest.mark.timeout(900)
@pytest.mark.parametrize("fmt", ["bf16", "f16", "f32"])
def test_c3_dmumu_within_one_percent_of_fp64(c3_runs, fmt):
    curves, lags = c3_runs
    rel_b, z_b, ok_b = compare(curves[fmt], curves["fp64"], "bin")
    rel_l, z_l, ok_l = compare(curves[fmt], curves["fp64"], "lag")
    print(f"\n{fmt}: |mu| bins: median rel err {np.median(rel_b):.2e}, max {rel_b.max():.2e}, "
          f"max err/sigma {z_b.max():.2f}; per bin: " + " ".join(f"{v:.1e}" for v in rel_b)
          + "; per bin err/sigma: " + " ".join(f"{v:.2f}" for v in z_b))
    print(f"{fmt}: lags ({len(lags)}, 1..{lags[-1]} samples): median rel err {np.median(rel_l):.2e}, "
          f"max {rel_l.max():.2e}, max err/sigma {z_l.max():.2f}")
    assert len(rel_b) == 20
    assert ok_b.all(), f"bins failing: {np.flatnonzero(~ok_b)}"
    assert ok_l.all(), f"lags failing: {np.flatnonzero(~ok_l)}"
