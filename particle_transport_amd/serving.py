"""Continuous-batching transport-campaign server over pt_serve_*.

``Server`` wraps the C-ABI server: fields stay resident, requests are admitted
FIFO against an HBM budget for particle state plus cadence cache, and every
admitted request advances in one fused launch per tick.  ``run_stream``
replays a timed request stream (Poisson arrivals), computes each finished
request's product (D_mumu from its mu cache, or D_pp from its |p| cache) on a
second context so it overlaps the next tick, and reports campaign latency
percentiles and aggregate particle-steps/s.

The campaign stream is the serving configuration of the benchmark plan:
field id uniform over the resident fields, energy log-uniform over
[1 GeV, 100 TeV], particle count log-uniform over [1e4, 1e6], duration uniform
over [10, 500] T_g0 at 200 steps/T_g0, isotropic or fixed-mu injection (mu0
uniform on [-1, 1]), output mu cache + D_mumu (64 lags x 20 bins) or D_pp.
"""
from __future__ import annotations

import ctypes
import math
import time
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import check, ptr
from .injection import C_LIGHT, M_P, Q_E, energy_to_gamma

QUEUED, RUNNING, DONE, RELEASED = 0, 1, 2, 3
OUT_NONE, OUT_MU, OUT_PMAG, OUT_DPREL = 0, 2, 3, 4


def cache_dtype(out_kind):
    """Element type of a request's cadence cache."""
    return np.float64 if out_kind == OUT_PMAG else np.float32


class ServeRequest(ctypes.Structure):
    _fields_ = [("field_id", ctypes.c_int32), ("out_kind", ctypes.c_int32), ("n", ctypes.c_int64),
                ("dt_s", ctypes.c_double), ("charge_C", ctypes.c_double), ("mass_kg", ctypes.c_double),
                ("total_steps", ctypes.c_int64), ("save_stride", ctypes.c_int64), ("periodic", ctypes.c_int32),
                ("inject_fixed_mu", ctypes.c_int32), ("mu0", ctypes.c_double), ("p_magnitude", ctypes.c_double),
                ("x", ctypes.POINTER(ctypes.c_double) * 3), ("p", ctypes.POINTER(ctypes.c_double) * 3),
                ("phase", ctypes.POINTER(ctypes.c_double)), ("inject_on_device", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("inject_first_index", ctypes.c_int64),
                ("inject", _lib.InjectParams), ("out_p_ref", ctypes.c_double)]


def _bind(L):
    if getattr(L, "_serve_bound", False):
        return L
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    i64p = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "pt_serve_create": [vp, i64, i32, i32, ctypes.POINTER(vp)],
        "pt_serve_destroy": [vp],
        "pt_serve_add_field": [vp, vp, ctypes.POINTER(i32)],
        "pt_serve_submit": [vp, ctypes.POINTER(ServeRequest), i64p],
        "pt_serve_step": [vp, i64p, i64p, i64p],
        "pt_serve_drain": [vp],
        "pt_serve_take_done": [vp, i64p, i64, i64p],
        "pt_serve_poll": [vp, i64, ctypes.POINTER(i32)],
        "pt_serve_fetch": [vp, i64, ctypes.POINTER(vp), i64p, i64p, i64p],
        "pt_serve_copy_cache": [vp, i64, vp, i64],
        "pt_serve_release": [vp, i64],
        "pt_serve_stats": [vp, i64p, i64p, ctypes.POINTER(ctypes.c_double)],
        "pt_dpp_accumulate_pmag": [vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int, i64, i64, i64, i64p, i64p,
                                   i64, ctypes.c_double, i64p, _lib._dp, _lib._dp, _lib._dp, _lib._dp],
        "pt_dpp_accumulate_dprel": [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, i64, i64, i64, i64p, i64p,
                                    i64, ctypes.c_double, ctypes.c_double, i64p, _lib._dp, _lib._dp, _lib._dp,
                                    _lib._dp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    L._serve_bound = True
    return L


class Server:
    def __init__(self, ctx, budget_bytes: int, steps_per_tick: int = 200, max_active: int = 1024):
        self.L = _bind(_lib.load())
        h = ctypes.c_void_p()
        check(self.L.pt_serve_create(ctx.h, int(budget_bytes), int(steps_per_tick), int(max_active),
                                     ctypes.byref(h)))
        self.h = h
        self.ctx = ctx

    def add_field(self, field) -> int:
        fid = ctypes.c_int32()
        check(self.L.pt_serve_add_field(self.h, field.h, ctypes.byref(fid)))
        return fid.value

    def submit(self, field_id, x=None, p=None, *, dt, total_steps, stride, out_kind=OUT_MU, charge=Q_E, mass=M_P,
               periodic=True, mu0=None, phase=None, p_magnitude=None, n=None, inject=None, first_index=0,
               p_ref=None) -> int:
        """Host state: x (and p, or mu0 + phase + p_magnitude for fixed-mu).
        Device state: n particles and `inject` (a _lib.InjectParams), drawn at
        admission as pt_inject draws particles first_index .. first_index+n-1."""
        r = ServeRequest()
        r.field_id, r.out_kind = int(field_id), int(out_kind)
        r.dt_s, r.charge_C, r.mass_kg = float(dt), float(charge), float(mass)
        r.total_steps, r.save_stride, r.periodic = int(total_steps), int(stride), int(bool(periodic))
        if out_kind == OUT_DPREL:
            r.out_p_ref = float(p_ref)
        if inject is not None:
            r.n, r.inject_on_device, r.inject_first_index, r.inject = int(n), 1, int(first_index), inject
            rid = ctypes.c_int64()
            check(self.L.pt_serve_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
            return rid.value
        n = len(x[0])
        r.n = n
        keep = [np.ascontiguousarray(v, dtype=np.float64) for v in x]
        for c in range(3):
            r.x[c] = keep[c].ctypes.data_as(_lib._dp)
        if mu0 is not None:
            ph = np.ascontiguousarray(phase, dtype=np.float64)
            keep.append(ph)
            r.inject_fixed_mu, r.mu0, r.p_magnitude = 1, float(mu0), float(p_magnitude)
            r.phase = ph.ctypes.data_as(_lib._dp)
        else:
            pk = [np.ascontiguousarray(v, dtype=np.float64) for v in p]
            keep += pk
            for c in range(3):
                r.p[c] = pk[c].ctypes.data_as(_lib._dp)
        rid = ctypes.c_int64()
        check(self.L.pt_serve_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    def step(self):
        a, q, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self.L.pt_serve_step(self.h, ctypes.byref(a), ctypes.byref(q), ctypes.byref(d)))
        return a.value, q.value, d.value

    def drain(self):
        check(self.L.pt_serve_drain(self.h))

    def take_done(self, capacity=4096):
        """Ids of requests that became DONE since the previous call."""
        buf = (ctypes.c_int64 * capacity)()
        cnt = ctypes.c_int64()
        check(self.L.pt_serve_take_done(self.h, buf, capacity, ctypes.byref(cnt)))
        return list(buf[:cnt.value])

    def poll(self, rid) -> int:
        s = ctypes.c_int32()
        check(self.L.pt_serve_poll(self.h, int(rid), ctypes.byref(s)))
        return s.value

    def fetch(self, rid):
        """(device pointer of the [save][n] cache, nsave, n, alive_final)."""
        dp, ns, n, al = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self.L.pt_serve_fetch(self.h, int(rid), ctypes.byref(dp), ctypes.byref(ns), ctypes.byref(n),
                                    ctypes.byref(al)))
        return dp.value, ns.value, n.value, al.value

    def copy_cache(self, rid, dtype):
        _, ns, n, _ = self.fetch(rid)
        out = np.empty((ns, n), dtype=dtype)
        check(self.L.pt_serve_copy_cache(self.h, int(rid), ptr(out), out.nbytes))
        return out

    def release(self, rid):
        check(self.L.pt_serve_release(self.h, int(rid)))

    def stats(self):
        b, ps, ms = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        check(self.L.pt_serve_stats(self.h, ctypes.byref(b), ctypes.byref(ps), ctypes.byref(ms)))
        return SimpleNamespace(bytes_reserved=b.value, particle_steps=ps.value, kernel_ms=ms.value)

    def close(self):
        if self.h:
            check(self.L.pt_serve_destroy(self.h))
            self.h = None


def dmumu_on_device(ctx, cache_ptr, nsave, n, lag_steps, edges, bin_abs=True, start_mode="sliding"):
    """D_mumu accumulators from a device mu cache [save][n] (float32)."""
    from .dmumu import new_accumulators
    acc = new_accumulators(len(lag_steps), len(edges) - 1)
    lags = np.ascontiguousarray(lag_steps, dtype=np.int64)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    L = _lib.load()
    check(L.pt_dmumu_accumulate(ctx.h, ctypes.cast(ctypes.c_void_p(cache_ptr), _lib._fp), 1, nsave, n, n, 1, None,
                                ptr(lags, _lib._i64p), len(lags), ptr(e, _lib._dp), len(e) - 1,
                                {"sliding": 0, "injection": 1}[start_mode], int(bin_abs),
                                ptr(acc.particle_counts, _lib._i64p), ptr(acc.particle_means, _lib._dp),
                                ptr(acc.particle_m2s, _lib._dp), ptr(acc.pair_sum_squares, _lib._dp),
                                ptr(acc.pair_counts, _lib._i64p), ptr(acc.dmumu_counts, _lib._i64p),
                                ptr(acc.dmumu_sum_delta, _lib._dp), ptr(acc.dmumu_sum_delta2, _lib._dp)))
    return acc


def dpp_on_device(ctx, cache_ptr, nsave, n, lag_steps, p0):
    """D_pp accumulators from a device |p| cache [save][n] (float64)."""
    from .dpp import new_accumulators
    acc = new_accumulators(len(lag_steps))
    lags = np.ascontiguousarray(lag_steps, dtype=np.int64)
    L = _bind(_lib.load())
    check(L.pt_dpp_accumulate_pmag(ctx.h, ctypes.cast(ctypes.c_void_p(cache_ptr), _lib._dp), 1, nsave, n, n, None,
                                   ptr(lags, _lib._i64p), len(lags), float(p0), ptr(acc.counts, _lib._i64p),
                                   ptr(acc.sum_delta_p, _lib._dp), ptr(acc.sum_delta_p2, _lib._dp),
                                   ptr(acc.sum_delta_p_norm, _lib._dp), ptr(acc.sum_delta_p_norm2, _lib._dp)))
    return acc


def dpp_dprel_on_device(ctx, cache_ptr, nsave, n, lag_steps, p_ref, p0):
    """D_pp accumulators from a device relative cache |p|/p_ref - 1 [save][n] (float32)."""
    from .dpp import new_accumulators
    acc = new_accumulators(len(lag_steps))
    lags = np.ascontiguousarray(lag_steps, dtype=np.int64)
    L = _bind(_lib.load())
    check(L.pt_dpp_accumulate_dprel(ctx.h, ctypes.cast(ctypes.c_void_p(cache_ptr), _lib._fp), 1, nsave, n, n, None,
                                    ptr(lags, _lib._i64p), len(lags), float(p_ref), float(p0),
                                    ptr(acc.counts, _lib._i64p), ptr(acc.sum_delta_p, _lib._dp),
                                    ptr(acc.sum_delta_p2, _lib._dp), ptr(acc.sum_delta_p_norm, _lib._dp),
                                    ptr(acc.sum_delta_p_norm2, _lib._dp)))
    return acc


def log_lag_steps(nsave, count):
    """Up to ``count`` log-spaced cadence offsets in [1, nsave-1], deduplicated."""
    if nsave < 2:
        return np.zeros(0, dtype=np.int64)
    steps = np.unique(np.round(np.geomspace(1, nsave - 1, count)).astype(np.int64))
    return steps[steps >= 1]


def make_stream(n_requests, n_fields, B0s, *, rate_hz=20.0, seed=7, e_range=(1.0, 1e5), n_range=(1e4, 1e6),
                dur_range=(10.0, 500.0), steps_per_gp=200, stride=20, box_length_m=None):
    """Request descriptors of the serving stream (host arrays built lazily)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests))
    reqs = []
    for i in range(n_requests):
        fid = int(rng.integers(n_fields))
        E = float(np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]))))
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        dur = float(rng.uniform(*dur_range))
        fixed = bool(rng.random() < 0.5)
        mu0 = float(rng.uniform(-1, 1)) if fixed else None
        kind = OUT_MU if rng.random() < 0.5 else OUT_DPREL  # D_pp from the relative |p| cache
        gamma = float(energy_to_gamma(E))
        Omega0 = Q_E * B0s[fid] / (gamma * M_P)
        dt = (2 * math.pi / steps_per_gp) / Omega0
        total = int(math.floor(dur * steps_per_gp)) + 1
        reqs.append(SimpleNamespace(i=i, arrival=float(t[i]), field_id=fid, energy_gev=E, n=n, duration_gp=dur,
                                    mu0=mu0, out_kind=kind, dt=dt, total_steps=total, stride=stride, gamma=gamma,
                                    seed=int(rng.integers(1 << 31))))
    return reqs


def request_inject_params(req):
    """The request's device injection: uniform positions in the box, fixed-mu
    (req.mu0) or isotropic directions, |p| = gamma m_p v (init_particles,
    phase_space_gpu_runner.jl:196-245), Philox4x64-10 keyed by req.seed."""
    ip = _lib.InjectParams()
    ip.seed = int(req.seed) & ((1 << 64) - 1)
    ip.fixed_position = 0
    ip.fixed_mu = int(req.mu0 is not None)
    ip.mu0 = float(req.mu0) if req.mu0 is not None else 0.0
    ip.p_magnitude = req.gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / req.gamma ** 2)
    return ip


def request_state(req, L):
    """Initial positions (uniform) and momenta/phases of one request, drawn
    on the host (numpy) for submissions that pass host arrays."""
    rng = np.random.default_rng(req.seed)
    x = rng.random((3, req.n)) * L
    pmag = req.gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / req.gamma ** 2)
    if req.mu0 is not None:
        return x, None, 2 * math.pi * rng.random(req.n), pmag
    z = 2 * rng.random(req.n) - 1
    phi = 2 * math.pi * rng.random(req.n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    return x, pmag * np.stack([r * np.cos(phi), r * np.sin(phi), z]), None, pmag


def submit_request(server, req, L, on_device=True):
    """Submit one stream request.  on_device: its initial state is drawn at
    admission (request_inject_params), so the host does no O(n) work per
    request; otherwise request_state draws it on the host."""
    pref = req.gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / req.gamma ** 2)  # the injected |p|
    if on_device:
        return server.submit(req.field_id, n=req.n, inject=request_inject_params(req), dt=req.dt,
                             total_steps=req.total_steps, stride=req.stride, out_kind=req.out_kind, p_ref=pref)
    x, p, phase, pmag = request_state(req, L)
    return server.submit(req.field_id, x, p, dt=req.dt, total_steps=req.total_steps, stride=req.stride,
                         out_kind=req.out_kind, mu0=req.mu0, phase=phase, p_magnitude=pmag, p_ref=pref)


def request_product(stats_ctx, cache_ptr, nsave, n, req):
    """The request's product from its device cache: D_mumu (64 log lags x 20
    |mu| bins) or D_pp."""
    lags = log_lag_steps(nsave, 64)
    if len(lags) == 0:
        return None
    if req.out_kind == OUT_MU:
        return dmumu_on_device(stats_ctx, cache_ptr, nsave, n, lags, np.linspace(0.0, 1.0, 21))
    pmag = req.gamma * M_P * C_LIGHT * math.sqrt(1 - 1 / req.gamma ** 2)  # p0: every particle starts at it
    if req.out_kind == OUT_DPREL:
        return dpp_dprel_on_device(stats_ctx, cache_ptr, nsave, n, lags, pmag, pmag)
    return dpp_on_device(stats_ctx, cache_ptr, nsave, n, lags, pmag)


def finish_request(server, stats_ctx, rid, req):
    """request_product of a finished request of `server`."""
    cp, nsave, n, alive = server.fetch(rid)
    return request_product(stats_ctx, cp, nsave, n, req)


def run_stream(push_ctx, stats_ctx, fields, reqs, *, budget_bytes, steps_per_tick=200, max_active=1024,
               box_length_m, time_scale=1.0, keep_results=False, keep_products=False, log=None):
    """Replay the stream in real time (arrival times x time_scale).  The loop
    keeps the server stepping; each finished request's product is computed on
    `stats_ctx` by a worker thread (the bindings release the GIL), and the
    request is released once its product is done.  Latency = arrival to
    product.

    keep_results: True or a set of request indices whose (product, host copy
    of the cache) are returned in `results`; keep_products: return every
    request's product in `products`.  `peak_reserved_bytes` is the largest
    budget the server held at any step."""
    from concurrent.futures import ThreadPoolExecutor

    srv = Server(push_ctx, budget_bytes, steps_per_tick, max_active)
    fids = [srv.add_field(f) for f in fields]
    busy = [0.0]
    failed = []

    def timed_product(*a):
        t = time.perf_counter()
        try:
            return request_product(*a)
        finally:
            busy[0] += time.perf_counter() - t

    assert fids == list(range(len(fields)))
    pending = sorted(reqs, key=lambda r: r.arrival)
    outstanding = {}
    finishing = {}
    results = {}
    products = {}
    peak = [0]
    lat = []
    pool = ThreadPoolExecutor(max_workers=1)
    t0 = time.perf_counter()
    k = 0
    last_log = 0.0
    try:
        while k < len(pending) or outstanding or finishing:
            now = time.perf_counter() - t0
            while k < len(pending) and pending[k].arrival * time_scale <= now:
                rq = pending[k]
                outstanding[submit_request(srv, rq, box_length_m)] = rq
                k += 1
            running, queued, done = srv.step()
            peak[0] = max(peak[0], srv.stats().bytes_reserved)
            for rid in srv.take_done():
                rq = outstanding.pop(rid)
                try:
                    cp, nsave, n, _ = srv.fetch(rid)
                except _lib.PTError as e:  # e.g. fixed-mu injection where the local B is zero
                    failed.append((rq.i, str(e)))
                    srv.release(rid)
                    continue
                finishing[rid] = (pool.submit(timed_product, stats_ctx, cp, nsave, n, rq), rq)
            for rid in [r for r, (f, _) in finishing.items() if f.done()]:
                fut, rq = finishing.pop(rid)
                res = fut.result()
                lat.append(time.perf_counter() - t0 - rq.arrival * time_scale)
                if keep_results is True or (keep_results and rq.i in keep_results):
                    results[rq.i] = (res, srv.copy_cache(rid, cache_dtype(rq.out_kind)))
                if keep_products:
                    products[rq.i] = res
                srv.release(rid)
            if log and now - last_log >= 1.0:
                last_log = now
                log(f"[serve] t={now:.1f}s submitted={k} running={running} queued={queued} done={done}")
            if running == 0 and queued == 0 and not outstanding:
                if finishing:
                    time.sleep(0.0005)
                elif k < len(pending):
                    time.sleep(max(0.0, pending[k].arrival * time_scale - (time.perf_counter() - t0)))
        srv.drain()
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        pool.shutdown(wait=True)
        srv.close()
    lat = np.asarray(lat)
    return SimpleNamespace(latency_p50_s=float(np.percentile(lat, 50)), latency_p99_s=float(np.percentile(lat, 99)),
                           requests=len(lat), wall_s=wall, particle_steps=st.particle_steps,
                           particle_steps_per_s=st.particle_steps / wall, kernel_ms=st.kernel_ms,
                           product_busy_s=busy[0], failed=failed, results=results, products=products,
                           latencies=lat, peak_reserved_bytes=int(peak[0]))
