"""Cell tables (pt_field_cell_table): the production pushers read a cell's
precomputed coefficients instead of rebuilding them from the stored nodes.
The table is an acceleration structure only, so every output must be bit
for bit the same with and without it: trajectories, mu caches and alive
counts of the single-campaign pusher on every storage it serves, with
periodic and open boundaries, at an energy where a wave has lanes changing
cell almost every push (the reload path runs all the time), and the serving
tick's caches."""
import math

import numpy as np
import pytest

from particle_transport_amd import fields, serving
from particle_transport_amd.injection import C_LIGHT, M_P, Q_E, energy_to_gamma
from particle_transport_amd.runner import Ensemble, Field
from particle_transport_amd._lib import PTError

pytestmark = pytest.mark.gpu

L_BOX = fields.box_length_m(10.0)


def make_case(n_nodes=33, n=1000, energy_gev=1e5, seed=31, steps_per_gp=50):
    raw = fields.generate_turbulence(n_nodes, seed=seed, kmax=8, dB_over_B0=1.0)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    rng = np.random.default_rng(seed + 1)
    x = rng.random((3, n)) * L_BOX
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    gamma = float(energy_to_gamma(energy_gev))
    p = d * gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / gamma ** 2)
    return F, x, p, gamma, steps_per_gp


def run(ctx, F, x, p, gamma, S, storage, table, periodic, out_kind):
    fld = Field(ctx, F, L_BOX, storage, cell_table=table)
    dt = (2 * math.pi / S) / (Q_E * fld.B0 / (gamma * M_P))
    ens = Ensemble(ctx, fld, x, p, pusher="fast", periodic=periodic)
    r = ens.advance(dt=dt, first_step=1, n_steps=700, total_steps=700, stride=3, out_kind=out_kind,
                    steps_per_launch=128)
    xs, ps, alive = ens.state()
    ens.close()
    fld.close()
    return r, xs, ps, alive


@pytest.mark.parametrize("storage", ["f32_planar", "f32_brick", "bf16_brick", "f16_brick"])
@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "open"])
def test_cell_table_gives_bit_identical_trajectories(ctx, storage, periodic):
    F, x, p, gamma, S = make_case()
    a = run(ctx, F, x, p, gamma, S, storage, False, periodic, "phase")
    b = run(ctx, F, x, p, gamma, S, storage, True, periodic, "phase")
    assert np.array_equal(a[0].alive_counts, b[0].alive_counts)
    assert np.array_equal(a[0].out_a, b[0].out_a, equal_nan=True)
    assert np.array_equal(a[0].out_b, b[0].out_b, equal_nan=True)
    for u, v in zip(a[1:], b[1:]):
        assert np.array_equal(u, v, equal_nan=True)
    if not periodic:
        assert a[0].alive_counts[-1] < len(x[0])  # the open box lost particles through its faces
    # the case is reload-heavy: between saves (3 pushes) several percent of
    # the particles change cell, so a wave of 64 reloads most pushes
    cells = np.floor(a[0].out_a.reshape(-1, 3, len(x[0])) / (L_BOX / 32))
    changed = np.any(np.diff(cells, axis=0) != 0, axis=1)
    ok = np.all(np.isfinite(cells[1:]), axis=1) & np.all(np.isfinite(cells[:-1]), axis=1)
    assert changed[ok].mean() > 0.02


def test_cell_table_mu_cache_and_lower_energy(ctx):
    F, x, p, gamma, S = make_case(energy_gev=100.0, steps_per_gp=200)
    a = run(ctx, F, x, p, gamma, S, "bf16_brick", False, True, "mu")
    b = run(ctx, F, x, p, gamma, S, "bf16_brick", True, True, "mu")
    assert np.array_equal(a[0].out_a, b[0].out_a, equal_nan=True)
    assert np.isfinite(a[0].out_a).all()


def test_cell_table_can_be_freed_and_rebuilt(ctx):
    F, x, p, gamma, S = make_case(n=256)
    fld = Field(ctx, F, L_BOX, "f16_brick", cell_table=True)
    fld.set_cell_table(False)
    fld.set_cell_table(True)
    fld.set_cell_table(True)  # already present: no-op
    fld.close()
    f64 = Field(ctx, F, L_BOX, "f64_planar")
    with pytest.raises(PTError):
        f64.set_cell_table(True)
    f64.close()


def test_serving_with_cell_tables_equals_without(ctx):
    raw = fields.generate_turbulence(33, seed=21, kmax=8, dB_over_B0=1.0)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    outs = []
    for table in (False, True):
        fld = Field(ctx, F, L_BOX, "bf16_brick", cell_table=table)
        B0 = fld.B0
        srv = serving.Server(ctx, 1 << 32, 64)
        srv.add_field(fld)
        from types import SimpleNamespace
        rids = []
        for i, (E, mu0) in enumerate(((1e4, None), (1e5, 0.3), (30.0, None))):
            gamma = float(energy_to_gamma(E))
            dt = (2 * math.pi / 100) / (Q_E * B0 / (gamma * M_P))
            r = SimpleNamespace(i=i, arrival=0.0, field_id=0, energy_gev=E, n=900 + 300 * i, mu0=mu0,
                                out_kind=serving.OUT_MU, dt=dt, total_steps=301, stride=10, gamma=gamma, seed=7 + i)
            rids.append((serving.submit_request(srv, r, L_BOX, on_device=True), r))
        res = []
        while any(srv.poll(rid) != serving.DONE for rid, _ in rids):
            srv.step()
        for rid, r in rids:
            res.append(srv.copy_cache(rid, serving.cache_dtype(r.out_kind)))
            srv.release(rid)
        srv.close()
        fld.close()
        outs.append(res)
    for u, v in zip(*outs):
        assert np.array_equal(u, v, equal_nan=True)


def test_server_refuses_fields_that_differ_in_tables(ctx):
    raw = fields.generate_turbulence(17, seed=5, kmax=4, dB_over_B0=1.0)["total"]
    F = [np.asfortranarray(a) for a in fields.to_si(raw)]
    a = Field(ctx, F, L_BOX, "bf16_brick", cell_table=True)
    b = Field(ctx, F, L_BOX, "bf16_brick")
    srv = serving.Server(ctx, 1 << 30, 64)
    srv.add_field(a)
    with pytest.raises(PTError):
        srv.add_field(b)
    srv.close()
    a.close()
    b.close()
