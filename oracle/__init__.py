"""CPU oracle for the trajectory and lag-statistics path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker.  The product path
(``particle_transport_amd``) never imports it.

The arithmetic lives in ``pt_oracle.c`` (plain C, FP64, ``-ffp-contract=off``),
a restatement of the reference's Julia functions in their operation order;
this module is a thin numpy/ctypes wrapper.  Build with ``make -C oracle``.

Parity pinning of each function (what independent check pins it):
  - ``trilinear``       exact on trilinear polynomials and at grid nodes.
  - ``boris_push``      |p| conserved and rotation angle 2*atan(|t|) per step in
                        uniform B; E x B drift in crossed uniform fields.
  - ``run``             the two above, plus the wrap/kill rules and the cadence
                        rule 1:stride:nsteps checked on their own.
  - ``mu_from_pB``      closed-form vectors (parallel, perpendicular, 45 degrees,
                        zero norm, NaN).
  - ``dmumu``           pure-Python brute force on hand-built series, Welford
                        against two-pass variance, injection == sliding at nsave=2.
  - ``dpp``             pure-Python brute force on hand-built momentum series.
  - ``energy_snapshots`` the relativistic closed form E = (gamma-1) m c^2.
  - ``mean_abs_b``      the reference's own known answer (3,4,0) -> 5.
  - ``philox4x64_10``   numpy.random.Philox (independent implementation) on
                        arbitrary keys and counters.
  - ``inject``          positions equal numpy Philox doubles mapped onto the
                        box; fixed-mu gives p.b/(|p||b|) == mu0 and |p| ==
                        pscale; isotropic moments <u_z> = 0, <u_z^2> = 1/3.

Parity unpinned (nothing available here fixes these bits):
  - Julia's MersenneTwister streams (the reference's injection draws and its
    random / block-random particle selection): the parity tests hand both
    sides identical initial arrays and index sets instead.
  - the bit pattern of Julia's pairwise ``mean`` behind ``mean_abs_b``: the
    known answer pins the value, other fields are compared at 1e-15.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

C_LIGHT = 2.99792458e8
Q_E = 1.602176634e-19
M_P = 1.67262192369e-27

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64 = ctypes.c_int64
_d = ctypes.c_double


class Geom(ctypes.Structure):
    _fields_ = [("nx", _i64), ("ny", _i64), ("nz", _i64),
                ("xmin", _d), ("ymin", _d), ("zmin", _d),
                ("xmax", _d), ("ymax", _d), ("zmax", _d),
                ("dx", _d), ("dy", _d), ("dz", _d)]


def build() -> str:
    """Compile liboracle.so in place (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_trilinear.restype = _d
        L.oracle_trilinear.argtypes = [_dp, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i64, _i64, _i64]
        L.oracle_boris_push.restype = None
        L.oracle_boris_push.argtypes = [_d] * 13 + [_dp]
        L.oracle_wrap.restype = _d
        L.oracle_wrap.argtypes = [_d, _d, _d]
        L.oracle_mu_from_pB.restype = _d
        L.oracle_mu_from_pB.argtypes = [_d] * 6
        L.oracle_run.restype = _i64
        L.oracle_run.argtypes = [ctypes.POINTER(_dp), ctypes.POINTER(_dp), _u8p, ctypes.POINTER(_dp),
                                 ctypes.POINTER(Geom), _i64, _d, _i64, _i64, _i64, _i64, ctypes.c_int,
                                 ctypes.c_int, _dp, _dp, _dp, _i64p]
        L.oracle_reconstruct_mu_chunk.restype = None
        L.oracle_reconstruct_mu_chunk.argtypes = [_dp, _dp, ctypes.POINTER(_dp), ctypes.POINTER(Geom),
                                                  _i64, _i64, _dp]
        L.oracle_mu_bin_index.restype = _i64
        L.oracle_mu_bin_index.argtypes = [_d, _dp, _i64]
        L.oracle_combine_welford.restype = None
        L.oracle_combine_welford.argtypes = [_i64p, _dp, _dp, _i64, _i64, _d, _d]
        L.oracle_dmumu.restype = None
        L.oracle_dmumu.argtypes = [_fp, _i64, _i64, _i64p, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int,
                                   _i64p, _dp, _dp, _dp, _i64p, _i64p, _dp, _dp]
        L.oracle_dpp.restype = None
        L.oracle_dpp.argtypes = [_fp, _i64, _i64, _i64p, _i64, _d, _i64p, _dp, _dp, _dp, _dp]
        L.oracle_kinetic_energy_gev.restype = _d
        L.oracle_kinetic_energy_gev.argtypes = [_d]
        L.oracle_energy_snapshots.restype = None
        L.oracle_energy_snapshots.argtypes = [_fp, _i64, _i64p, _i64, _dp, _i64]
        L.oracle_mean_abs_b.restype = _d
        L.oracle_mean_abs_b.argtypes = [_dp, _dp, _dp, _i64]
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_philox4x64_10.restype = None
        L.oracle_philox4x64_10.argtypes = [_u64p, _u64p, _u64p]
        L.oracle_inject_draw.restype = _d
        L.oracle_inject_draw.argtypes = [ctypes.c_uint64, _i64, ctypes.c_int]
        L.oracle_inject.restype = ctypes.c_int
        L.oracle_inject.argtypes = [ctypes.c_uint64, _i64, _i64, ctypes.c_int, _dp, ctypes.c_int, _d, _d,
                                    ctypes.POINTER(_dp), ctypes.POINTER(Geom), _dp]
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def make_geom(nx, ny, nz, box_length_m):
    """Geometry as the reference builds it: range(0, L, length=n) per axis
    (phase_space_gpu_runner.jl:125-130), xmin=x[1], xmax=x[end], dx=x[2]-x[1]."""
    def axis(n):
        # Julia range(0.0, L, length=n): elements are correctly rounded i*L/(n-1)
        # when computed exactly; x[2]-x[1] == x[2] since x[1] == 0.
        step1 = float(np_fraction_lerp(box_length_m, 1, n - 1))
        return 0.0, float(box_length_m), step1
    g = Geom()
    g.nx, g.ny, g.nz = nx, ny, nz
    g.xmin, g.xmax, g.dx = axis(nx)
    g.ymin, g.ymax, g.dy = axis(ny)
    g.zmin, g.zmax, g.dz = axis(nz)
    return g


def np_fraction_lerp(L, i, d):
    from fractions import Fraction
    return float(Fraction(L) * i / d)


def trilinear(arr, x, y, z, g):
    arr = np.asfortranarray(arr, dtype=np.float64)
    return lib().oracle_trilinear(_ptr(arr), x, y, z, g.xmin, g.ymin, g.zmin, g.dx, g.dy, g.dz,
                                  g.nx, g.ny, g.nz)


def boris_push(p, x, e, b, dt):
    out = np.zeros(6)
    lib().oracle_boris_push(*map(float, p), *map(float, x), *map(float, e), *map(float, b), float(dt),
                            _ptr(out))
    return out[:3].copy(), out[3:].copy()


def wrap(x, xmin, xmax):
    return lib().oracle_wrap(x, xmin, xmax)


def mu_from_pB(p, b):
    return lib().oracle_mu_from_pB(*map(float, p), *map(float, b))


def run(fields, g, x, p, dt, first_step, n_steps, total_steps, stride, periodic=True, out_kind=2,
        alive=None):
    """Run the reference step loop on the CPU.  fields: 6 Fortran-ordered
    (nx,ny,nz) float64 arrays (Bx,By,Bz in T, ux,uy,uz in m/s).  x, p: (3,n)
    float64, updated in place.  Returns dict(pos, mom, mu, alive_counts, nsaved)."""
    L = lib()
    n = x.shape[1]
    assert x.dtype == np.float64 and p.dtype == np.float64 and x.flags.c_contiguous and p.flags.c_contiguous
    F = [np.asfortranarray(f, dtype=np.float64) for f in fields]
    Fp = (_dp * 6)(*[_ptr(f) for f in F])
    xs = (_dp * 3)(*[x[c].ctypes.data_as(_dp) for c in range(3)])
    ps = (_dp * 3)(*[p[c].ctypes.data_as(_dp) for c in range(3)])
    if alive is None:
        alive = np.ones(n, dtype=np.uint8)
    nsave_max = sum(1 for k in range(n_steps) if (first_step + k - 1) % stride == 0)
    pos = np.full((max(nsave_max, 1), 3, n), np.nan) if out_kind == 1 else np.zeros(1)
    mom = np.full((max(nsave_max, 1), 3, n), np.nan) if out_kind == 1 else np.zeros(1)
    mu = np.full((max(nsave_max, 1), n), np.nan) if out_kind == 2 else np.zeros(1)
    counts = np.zeros(n_steps, dtype=np.int64)
    nsaved = L.oracle_run(xs, ps, _ptr(alive, _u8p), Fp, ctypes.byref(g), n, dt, first_step, n_steps,
                          total_steps, stride, int(periodic), out_kind, _ptr(pos), _ptr(mom), _ptr(mu),
                          _ptr(counts, _i64p))
    return dict(pos=pos[:nsaved] if out_kind == 1 else None, mom=mom[:nsaved] if out_kind == 1 else None,
                mu=mu[:nsaved] if out_kind == 2 else None, alive_counts=counts, nsaved=nsaved,
                alive=alive)


def reconstruct_mu_chunk(positions, momenta, Bxyz, g):
    """positions/momenta: C arrays (nsave,3,n) (Julia (n,3,nsave)); returns
    mu as C array (n, nsave) == Julia (nsave, n) column-major."""
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    momenta = np.ascontiguousarray(momenta, dtype=np.float64)
    nsave, _, n = positions.shape
    F = [np.asfortranarray(f, dtype=np.float64) for f in Bxyz]
    Fp = (_dp * 3)(*[_ptr(f) for f in F])
    mu = np.empty((n, nsave))
    lib().oracle_reconstruct_mu_chunk(_ptr(positions), _ptr(momenta), Fp, ctypes.byref(g), n, nsave,
                                      _ptr(mu))
    return mu


def mu_bin_index(v, edges):
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    return lib().oracle_mu_bin_index(float(v), _ptr(edges), len(edges))


def new_dmumu_accumulators(nlags, nbins):
    return dict(particle_counts=np.zeros(nlags, np.int64), particle_means=np.zeros(nlags),
                particle_m2s=np.zeros(nlags), pair_sum_squares=np.zeros(nlags),
                pair_counts=np.zeros(nlags, np.int64),
                dmumu_counts=np.zeros((nlags, nbins), np.int64),  # C (nlags,nbins) == Julia (nbins,nlags)
                dmumu_sum_delta=np.zeros((nlags, nbins)), dmumu_sum_delta2=np.zeros((nlags, nbins)))


def dmumu(mu_chunk, lag_steps, edges, injection, bin_abs, acc):
    """mu_chunk: C array (chunk, nsave) float32 (== Julia (nsave, chunk))."""
    mu_chunk = np.ascontiguousarray(mu_chunk, dtype=np.float32)
    chunk, nsave = mu_chunk.shape
    lag = np.ascontiguousarray(lag_steps, dtype=np.int64)
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    lib().oracle_dmumu(_ptr(mu_chunk, _fp), nsave, chunk, _ptr(lag, _i64p), len(lag), _ptr(edges),
                       len(edges), int(injection), int(bin_abs),
                       _ptr(acc["particle_counts"], _i64p), _ptr(acc["particle_means"]),
                       _ptr(acc["particle_m2s"]), _ptr(acc["pair_sum_squares"]),
                       _ptr(acc["pair_counts"], _i64p), _ptr(acc["dmumu_counts"], _i64p),
                       _ptr(acc["dmumu_sum_delta"]), _ptr(acc["dmumu_sum_delta2"]))
    return acc


def new_dpp_accumulators(nlags):
    return dict(counts=np.zeros(nlags, np.int64), sum_delta_p=np.zeros(nlags), sum_delta_p2=np.zeros(nlags),
                sum_delta_p_norm=np.zeros(nlags), sum_delta_p_norm2=np.zeros(nlags))


def dpp(momenta, lag_steps, p0, acc):
    """momenta: C array (nsave, 3, chunk) float32 (== Julia (chunk,3,nsave))."""
    momenta = np.ascontiguousarray(momenta, dtype=np.float32)
    nsave, _, chunk = momenta.shape
    lag = np.ascontiguousarray(lag_steps, dtype=np.int64)
    lib().oracle_dpp(_ptr(momenta, _fp), chunk, nsave, _ptr(lag, _i64p), len(lag), float(p0),
                     _ptr(acc["counts"], _i64p), _ptr(acc["sum_delta_p"]), _ptr(acc["sum_delta_p2"]),
                     _ptr(acc["sum_delta_p_norm"]), _ptr(acc["sum_delta_p_norm2"]))
    return acc


def kinetic_energy_gev(pm):
    return lib().oracle_kinetic_energy_gev(float(pm))


def energy_snapshots(momenta, snapshot_idx, out, col0):
    momenta = np.ascontiguousarray(momenta, dtype=np.float32)
    nsave, _, chunk = momenta.shape
    idx = np.ascontiguousarray(snapshot_idx, dtype=np.int64)
    assert out.flags.f_contiguous and out.dtype == np.float64
    lib().oracle_energy_snapshots(_ptr(momenta, _fp), chunk, _ptr(idx, _i64p), len(idx), _ptr(out), col0)
    return out


def mean_abs_b(bx, by, bz):
    bx = np.ascontiguousarray(np.ravel(bx, order="F"), dtype=np.float64)
    by = np.ascontiguousarray(np.ravel(by, order="F"), dtype=np.float64)
    bz = np.ascontiguousarray(np.ravel(bz, order="F"), dtype=np.float64)
    return lib().oracle_mean_abs_b(_ptr(bx), _ptr(by), _ptr(bz), bx.size)


def philox4x64_10(counter, key):
    """Philox4x64-10 block: four uint64 counter words, two key words."""
    c = (ctypes.c_uint64 * 4)(*[int(v) for v in counter])
    k = (ctypes.c_uint64 * 2)(*[int(v) for v in key])
    o = (ctypes.c_uint64 * 4)()
    lib().oracle_philox4x64_10(c, k, o)
    return [int(v) for v in o]


def inject_draw(seed, i, j):
    """Draw j of global particle i of the device injection, in [0, 1)."""
    return lib().oracle_inject_draw(int(seed), int(i), int(j))


def inject(seed, first, count, g, *, fixed_position=False, position=(0.0, 0.0, 0.0), fixed_mu=False, mu0=0.0,
           pscale=1.0, B=None):
    """init_particles with the device injection's draws for particles
    first..first+count-1.  B: three (nx,ny,nz) float64 arrays (fixed mu).
    Returns (out (6, count) float64, status)."""
    out = np.empty((6, count))
    pos = np.ascontiguousarray(position, dtype=np.float64)
    keep = [np.asfortranarray(b, dtype=np.float64) for b in B] if B is not None else [np.zeros(1)] * 3
    Bp = (_dp * 3)(*[_ptr(b) for b in keep])
    st = lib().oracle_inject(int(seed), int(first), int(count), int(bool(fixed_position)), _ptr(pos),
                             int(bool(fixed_mu)), float(mu0), float(pscale), Bp, ctypes.byref(g), _ptr(out))
    return out, st
