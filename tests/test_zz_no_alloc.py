"""pt_dmumu_accumulate_noalloc: the D_mumu pass with every work area in the
library's fixed device scratch array (no allocation), against
pt_dmumu_accumulate on the same device mu cache.  Both run the same kernels,
and the bin sums are exact integer (fixed-point) sums, so every accumulator
must be bitwise equal; the oracle pins both.  Two calls in a row, the second
smaller, check that the shared scratch is cleared between calls."""
import ctypes

import numpy as np
import pytest

import oracle
from particle_transport_amd import _lib, dmumu
from particle_transport_amd._lib import ptr

pytestmark = pytest.mark.gpu


def synthetic_mu(nsave, n, seed):
    """A mu cache [save][n] with slow drifts, NaN tails (deaths) and a few
    values on bin edges."""
    rng = np.random.default_rng(seed)
    mu = np.clip(np.cumsum(rng.normal(0, 0.05, (nsave, n)), axis=0) + rng.uniform(-1, 1, n), -1, 1)
    mu = mu.astype(np.float32)
    for j in rng.choice(n, n // 10, replace=False):
        mu[rng.integers(1, nsave):, j] = np.nan
    mu[rng.integers(0, nsave, 20), rng.integers(0, n, 20)] = 0.5
    return mu


def call(ctx, fn, mu_dev, nsave, n, idx, lags, edges, mode, bin_abs):
    acc = dmumu.new_accumulators(len(lags), len(edges) - 1)
    lags = np.ascontiguousarray(lags, dtype=np.int64)
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    idx_p = ptr(idx, _lib._i64p) if idx is not None else None
    nsel = len(idx) if idx is not None else n
    head = [ctx.h, ctypes.cast(ctypes.c_void_p(mu_dev.ptr), _lib._fp)]
    if fn == "pt_dmumu_accumulate":
        head.append(1)
    args = head + [nsave, nsel, n, 1, idx_p, ptr(lags, _lib._i64p), len(lags), ptr(edges, _lib._dp),
                   len(edges) - 1, mode, int(bin_abs),
                   ptr(acc.particle_counts, _lib._i64p), ptr(acc.particle_means, _lib._dp),
                   ptr(acc.particle_m2s, _lib._dp), ptr(acc.pair_sum_squares, _lib._dp),
                   ptr(acc.pair_counts, _lib._i64p), ptr(acc.dmumu_counts, _lib._i64p),
                   ptr(acc.dmumu_sum_delta, _lib._dp), ptr(acc.dmumu_sum_delta2, _lib._dp)]
    _lib.check(getattr(_lib.load(), fn)(*args))
    return acc


FIELDS = ("particle_counts", "particle_means", "particle_m2s", "pair_sum_squares", "pair_counts",
          "dmumu_counts", "dmumu_sum_delta", "dmumu_sum_delta2")


@pytest.mark.parametrize("mode,bin_abs", [(0, True), (1, False)])
def test_noalloc_equals_allocating_entry_point(ctx, mode, bin_abs):
    edges = np.linspace(0.0 if bin_abs else -1.0, 1.0, 21)
    # call 1: 3000 particles x 300 samples, every particle; call 2: a smaller
    # selection of a smaller cache, which must not see call 1's sums
    cases = [(300, 3000, None, [1, 2, 3, 5, 8, 13, 40, 120, 299]),
             (120, 700, np.arange(5, 700, 3, dtype=np.int64), [1, 4, 9, 30])]
    for k, (nsave, n, idx, lags) in enumerate(cases):
        mu = synthetic_mu(nsave, n, seed=k)
        buf = _lib.DeviceBuffer.from_host(ctx, mu)
        used = ctx.bytes_in_use()
        a = call(ctx, "pt_dmumu_accumulate_noalloc", buf, nsave, n, idx, lags, edges, mode, bin_abs)
        assert ctx.bytes_in_use() == used  # nothing allocated through the context
        b = call(ctx, "pt_dmumu_accumulate", buf, nsave, n, idx, lags, edges, mode, bin_abs)
        for f in FIELDS:
            assert np.array_equal(getattr(a, f), getattr(b, f)), (k, f)
        cols = mu if idx is None else mu[:, idx]
        o = oracle.new_dmumu_accumulators(len(lags), len(edges) - 1)
        oracle.dmumu(np.ascontiguousarray(cols.T), lags, edges, mode == 1, bin_abs, o)
        assert np.array_equal(a.dmumu_counts, o["dmumu_counts"])
        assert np.array_equal(a.pair_counts, o["pair_counts"])
        assert np.array_equal(a.particle_counts, o["particle_counts"])
        scale = np.maximum(np.abs(o["dmumu_sum_delta2"]), 1.0)
        assert np.max(np.abs(a.dmumu_sum_delta2 - o["dmumu_sum_delta2"]) / scale) <= 1e-11
        assert np.max(np.abs(a.dmumu_sum_delta - o["dmumu_sum_delta"]) / scale) <= 1e-11
        assert np.allclose(a.particle_means, o["particle_means"], rtol=1e-11, atol=0)
        buf.free()


def test_noalloc_rejects_calls_beyond_its_scratch(ctx):
    """A call whose work areas exceed the fixed 32 MiB is refused with
    PT_ERR_MEMORY before any kernel runs: 2^24 particles need 64 MiB for the
    per-particle finite-prefix bounds alone.  The device buffer still covers
    the whole extent the call describes (2 samples x 2^24 particles), so no
    kernel could read past it even if the check let the call through."""
    n_sel, nsave = 1 << 24, 2
    assert n_sel * 4 > 32 << 20
    buf = _lib.DeviceBuffer(ctx, (nsave, n_sel), np.float32)
    acc = dmumu.new_accumulators(1, 4)
    lags = np.array([1], dtype=np.int64)
    edges = np.linspace(0, 1, 5)
    rc = _lib.load().pt_dmumu_accumulate_noalloc(
        ctx.h, ctypes.cast(ctypes.c_void_p(buf.ptr), _lib._fp), nsave, n_sel, n_sel, 1, None,
        ptr(lags, _lib._i64p), 1, ptr(edges, _lib._dp), 4, 0, 1,
        ptr(acc.particle_counts, _lib._i64p), ptr(acc.particle_means, _lib._dp), ptr(acc.particle_m2s, _lib._dp),
        ptr(acc.pair_sum_squares, _lib._dp), ptr(acc.pair_counts, _lib._i64p), ptr(acc.dmumu_counts, _lib._i64p),
        ptr(acc.dmumu_sum_delta, _lib._dp), ptr(acc.dmumu_sum_delta2, _lib._dp))
    err = _lib.load().pt_last_error()
    buf.free()
    assert rc == _lib.PT_ERR_MEMORY, err
    assert b"scratch" in err
    assert acc.dmumu_counts.sum() == 0
