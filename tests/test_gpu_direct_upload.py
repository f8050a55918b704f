"""Cadence caches in page-locked host memory go to the device by strided
copies of the copy engine (pt_memcpy_2d), without a host read: the groups,
and the products built from them, equal those of the host-read path."""
import numpy as np
import pytest

from particle_transport_amd import _lib, dmumu, dpp, fields, runner

pytestmark = pytest.mark.gpu


def same(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return np.array_equal(x, y, equal_nan=x.dtype.kind in "fc")


def pinned_copy(a):
    p = _lib.PinnedArray(a.shape, a.dtype)
    p.array[...] = a
    return p.array


def test_strided_copy_equals_the_slice():
    rng = np.random.default_rng(3)
    src = rng.standard_normal((7, 3, 1001)).astype(np.float32)
    pin = pinned_copy(src)
    assert _lib.is_pinned(pin) and _lib.is_pinned(pin[2:5]) and not _lib.is_pinned(src)
    with runner.Context(0) as ctx:
        buf = _lib.DeviceBuffer(ctx, (7 * 3 * 1001,), np.float32)
        try:
            for lo, w in ((0, 1001), (0, 1), (17, 100), (1000, 1), (901, 100)):
                for s in (pin, src):  # page-locked, and pageable (staged by the runtime)
                    v = buf.view((7, 3, w))
                    v.upload_columns(s, lo)
                    assert np.array_equal(v.to_host(), src[:, :, lo:lo + w]), (lo, w)
            for lo, w in ((-1, 10), (992, 10)):
                with pytest.raises(ValueError):
                    buf.view((7, 3, w)).upload_columns(pin, lo)
            with pytest.raises(ValueError):
                buf.view((7, 2, 10)).upload_columns(pin, 0)
        finally:
            buf.free()


@pytest.mark.parametrize("selection", ["range", "random", "block_random"])
def test_device_groups_from_pinned_equal_host_groups(selection):
    rng = np.random.default_rng(11)
    n, nsave, chunk = 2000, 41, 64
    a = rng.standard_normal((nsave, 3, n)).astype(np.float32)
    pin = pinned_copy(a)
    cfg = dict(first_particle=5, n_particles_to_use=1500, particle_selection=selection, particle_seed=2,
               particle_block_size=100)
    pidx0 = dmumu.build_particle_indices(n, cfg) - 1
    with runner.Context(0) as ctx:
        for budget in (1 << 30, 40_000, 1):  # one span, several spans, per-chunk runs
            host = [(v.to_host(), [(c0, loc.copy()) for c0, loc in ch]) for (v,), ch in
                    dmumu.device_particle_groups(ctx, [a], pidx0, chunk, [np.float32], budget)]
            direct = [(v.to_host(), [(c0, loc.copy()) for c0, loc in ch]) for (v,), ch in
                      dmumu.device_particle_groups(ctx, [pin], pidx0, chunk, [np.float32], budget)]
            assert len(host) == len(direct) > 0
            for (hb, hc), (db, dc) in zip(host, direct):
                assert np.array_equal(hb, db)
                assert [c0 for c0, _ in hc] == [c0 for c0, _ in dc]
                assert all(np.array_equal(x, y) for (_, x), (_, y) in zip(hc, dc))
                for c0, loc in dc:
                    assert np.array_equal(db[..., loc], a[..., pidx0[c0:c0 + chunk]])


@pytest.mark.parametrize("selection", ["range", "random"])
def test_products_from_pinned_caches_equal_host_read(selection):
    """run_combined_full and run_dpp_full on page-locked caches (the runner's
    in-memory form) give the host-read path's products bit for bit."""
    raw = fields.generate_turbulence(17, seed=8, kmax=4, dB_over_B0=0.7)["total"]
    F = [np.asfortranarray(x, dtype=np.float32) for x in fields.to_si(raw)]
    L = fields.box_length_m(10.0)
    rng = np.random.default_rng(23)
    n, nsave = 700, 151
    pos = (rng.random((nsave, 3, n)) * L).astype(np.float32)
    d = rng.standard_normal((nsave, 3, n))
    mom = (d / np.linalg.norm(d, axis=1, keepdims=True) * 5.4e-18 * (1 + 0.01 * rng.random((nsave, 1, n))))
    mom = mom.astype(np.float32)
    t_s = 1.0e5 * np.arange(nsave)
    t_norm = t_s * 1e-6
    dcfg = dict(mu_min=0.0, mu_max=1.0, n_mu_bins=10, mu_bin_abs=True, dmumu_start_mode="sliding",
                first_particle=3, n_particles_to_use=600, particle_selection=selection, particle_seed=4,
                particle_chunk_size=64, gpu_batch_particles=256, lag_mode="uniform_samples",
                lag_range_policy="fixed", lag_boundary_policy="strict", lag_min_gyroperiods=None,
                lag_max_gyroperiods=None, min_lag_steps=1, max_lag_steps=None, n_lag_samples=9,
                min_count_per_cell=3, max_pair_visits_without_allow=5e9, allow_huge=False, n_energy_snapshots=4)
    ppos, pmom = pinned_copy(pos), pinned_copy(mom)
    with runner.Context(0) as ctx:
        fld = runner.Field(ctx, F, L, "f32_planar")
        try:
            want = dmumu.run_combined_full(dcfg, ctx, fld, pos, mom, t_s, t_norm)
            got = dmumu.run_combined_full(dcfg, ctx, fld, ppos, pmom, t_s, t_norm)
        finally:
            fld.close()
        want_p = dpp.run_dpp_full(dcfg, ctx, mom, t_s, t_norm)
        got_p = dpp.run_dpp_full(dcfg, ctx, pmom, t_s, t_norm)
    assert want["dmumu"]["count_pairs_full"].sum() > 0
    for grp in ("dmumu", "delta_mu2"):
        for k, v in want[grp].items():
            assert same(v, got[grp][k]), (grp, k)
    assert want_p["dpp"]["counts"].sum() > 0
    for grp in ("dpp", "energy_snapshots"):
        for k, v in want_p[grp].items():
            assert same(v, got_p[grp][k]), (grp, k)
