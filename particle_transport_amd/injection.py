"""Particle injection (host side): restates phase_space_gpu_runner.jl:51-93, 145-245.

``energy_to_gamma``/``energy_to_speed`` and ``init_particles`` keep the
reference's names, argument meaning, per-particle draw order and errors.  The
random stream is numpy's PCG64 seeded with ``cfg["seed"]``, not Julia's
MersenneTwister: the reference's exact stream cannot be reproduced here, so
parity runs hand identical initial arrays to both sides instead.

``init_particles_device`` and ``inject_params`` draw the same injection on the
device (``pt_inject`` / ``pt_ens_create_injected``) from a counter-based
Philox4x64-10 stream per particle, selected with ``injection_rng = "device"``.
"""
from __future__ import annotations

import math

import numpy as np

from .fields import PC_TO_M
from .time_units import PipelineError

C_LIGHT = 2.99792458e8
Q_E = 1.602176634e-19
M_P = 1.67262192369e-27
EV_TO_J = 1.602176634e-19
GEV_TO_J = 1e9 * EV_TO_J


def _T(dtype):
    return np.float64 if dtype in (np.float64, "Float64", "float64", float) else np.float32


def energy_to_gamma(E_GeV, dtype=np.float64):
    """phase_space_gpu_runner.jl:51-54, evaluated in precision T."""
    T = _T(dtype)
    kinetic = T(E_GeV * GEV_TO_J)
    return kinetic / (T(M_P) * T(C_LIGHT) ** 2) + T(1)


def energy_to_speed(E_GeV, dtype=np.float64):
    """phase_space_gpu_runner.jl:56-59."""
    T = _T(dtype)
    gamma = energy_to_gamma(E_GeV, T)
    return T(C_LIGHT) * np.sqrt(T(1) - T(1) / (gamma * gamma))


def _norm_mode(v):
    return str(v).strip().lower().replace("-", "_")


def injection_mode(cfg):
    m = _norm_mode(cfg.get("injection_mode", "isotropic"))
    if m not in ("isotropic", "fixed_mu"):
        raise PipelineError("injection_mode must be isotropic or fixed-mu")
    return m


def injection_position_mode(cfg):
    m = _norm_mode(cfg.get("injection_position_mode", "random"))
    if m not in ("random", "fixed"):
        raise PipelineError("injection_position_mode must be random or fixed")
    return m


def injection_position_unit(cfg):
    u = _norm_mode(cfg.get("injection_position_unit", "box_fraction"))
    if u not in ("box_fraction", "m", "pc"):
        raise PipelineError("injection_position_unit must be box-fraction, m, or pc")
    return u


def injection_position_tuple(cfg, T=np.float64):
    raw = cfg.get("injection_position", (0.5, 0.5, 0.5))
    if len(raw) != 3:
        raise PipelineError("injection_position must contain exactly three values")
    return tuple(T(v) for v in raw)


def fixed_injection_position(cfg, axes, T=np.float64):
    """phase_space_gpu_runner.jl:145-168."""
    px, py, pz = injection_position_tuple(cfg, T)
    unit = injection_position_unit(cfg)
    lo = [T(a["min"]) for a in axes]
    hi = [T(a["max"]) for a in axes]
    if unit == "box_fraction":
        pos = [lo[i] + (px, py, pz)[i] * (hi[i] - lo[i]) for i in range(3)]
    elif unit == "pc":
        pos = [v * T(PC_TO_M) for v in (px, py, pz)]
    else:
        pos = [px, py, pz]
    if not all(lo[i] <= pos[i] <= hi[i] for i in range(3)):
        raise PipelineError("fixed injection position is outside the simulation box")
    return tuple(pos)


def trilinear_sample_host(arr, x, y, z, axes):
    """trilinear_sample (phase_space_gpu_runner.jl:247-276) vectorised over
    particles, for the host-side fixed-mu injection.  arr is (nx,ny,nz)."""
    T = arr.dtype.type
    f = [(np.asarray(v, dtype=T) - T(a["min"])) / T(a["d"]) + T(1) for v, a in zip((x, y, z), axes)]
    idx = [np.clip(np.floor(fi), 1, a["n"] - 1).astype(np.int64) for fi, a in zip(f, axes)]
    t = [fi - ii.astype(T) for fi, ii in zip(f, idx)]
    i0, j0, k0 = (ii - 1 for ii in idx)
    tx, ty, tz = t
    g = lambda di, dj, dk: arr[i0 + di, j0 + dj, k0 + dk]  # noqa: E731
    c00 = g(0, 0, 0) * (1 - tx) + g(1, 0, 0) * tx
    c10 = g(0, 1, 0) * (1 - tx) + g(1, 1, 0) * tx
    c01 = g(0, 0, 1) * (1 - tx) + g(1, 0, 1) * tx
    c11 = g(0, 1, 1) * (1 - tx) + g(1, 1, 1) * tx
    c0 = c00 * (1 - ty) + c10 * ty
    c1 = c01 * (1 - ty) + c11 * ty
    return c0 * (1 - tz) + c1 * tz


def fixed_mu_unit_vectors(phi, bx, by, bz, mu0, T=np.float64):
    """fixed_mu_unit_vector (phase_space_gpu_runner.jl:170-194), vectorised;
    phi are the per-particle gyrophase draws 2*pi*rand."""
    bn = np.sqrt(bx * bx + by * by + bz * bz)
    if not np.all(np.isfinite(bn)) or np.any(bn == 0):
        raise PipelineError("cannot use fixed-mu injection where local B is zero or non-finite")
    bhx, bhy, bhz = bx / bn, by / bn, bz / bn
    small = np.abs(bhx) < T(0.9)
    e1x = np.where(small, T(0), -bhz)
    e1y = np.where(small, bhz, T(0))
    e1z = np.where(small, -bhy, bhx)
    e1n = np.sqrt(e1x * e1x + e1y * e1y + e1z * e1z)
    e1x, e1y, e1z = e1x / e1n, e1y / e1n, e1z / e1n
    e2x = bhy * e1z - bhz * e1y
    e2y = bhz * e1x - bhx * e1z
    e2z = bhx * e1y - bhy * e1x
    perp = np.sqrt(np.maximum(T(0), T(1) - mu0 * mu0))
    c, s = np.cos(phi), np.sin(phi)
    return (mu0 * bhx + perp * (c * e1x + s * e2x),
            mu0 * bhy + perp * (c * e1y + s * e2y),
            mu0 * bhz + perp * (c * e1z + s * e2z))


def init_particles(cfg, axes, gamma0, v0, dtype=np.float64, Bx=None, By=None, Bz=None):
    """init_particles (phase_space_gpu_runner.jl:196-245).

    Per particle, in the reference's order: three uniform position draws
    (unless the position is fixed), then either two direction draws
    (z = 2u-1, phi = 2 pi u) or, for fixed-mu, trilinear B at the position and
    one gyrophase draw.  Returns x1, x2, x3, p1, p2, p3 in precision T."""
    T = _T(dtype)
    rng = np.random.default_rng(int(cfg["seed"]))
    n = int(cfg["n_particles"])
    lo = [T(a["min"]) for a in axes]
    hi = [T(a["max"]) for a in axes]
    pscale = T(gamma0) * T(M_P) * T(v0)
    pos_mode = injection_position_mode(cfg)
    mom_mode = injection_mode(cfg)
    fixed_pos = fixed_injection_position(cfg, axes, T) if pos_mode == "fixed" else None
    mu0 = T(cfg.get("injection_mu0", 0.0))
    if mom_mode == "fixed_mu":
        if not (-1 <= mu0 <= 1):
            raise PipelineError("injection_mu0 must be in [-1, 1]")
        if Bx is None or By is None or Bz is None:
            raise PipelineError("fixed-mu injection requires magnetic field arrays")
    # Draws are consumed per particle in the reference's order; generating the
    # whole table at once with the same per-particle layout keeps that order.
    per = (3 if fixed_pos is None else 0) + (1 if mom_mode == "fixed_mu" else 2)
    u = rng.random((n, per))
    col = 0
    if fixed_pos is None:
        x = [(u[:, i].astype(T) * (hi[i] - lo[i]) + lo[i]).astype(T) for i in range(3)]
        col = 3
    else:
        x = [np.full(n, fixed_pos[i], dtype=T) for i in range(3)]
    if mom_mode == "fixed_mu":
        bx = trilinear_sample_host(np.asarray(Bx, dtype=T), *x, axes)
        by = trilinear_sample_host(np.asarray(By, dtype=T), *x, axes)
        bz = trilinear_sample_host(np.asarray(Bz, dtype=T), *x, axes)
        phi = (2 * math.pi * u[:, col]).astype(T)
        ux, uy, uz = fixed_mu_unit_vectors(phi, bx, by, bz, mu0, T)
    else:
        z = (2 * u[:, col] - 1).astype(T)
        phi = (2 * math.pi * u[:, col + 1]).astype(T)
        r = np.sqrt(np.maximum(T(0), T(1) - z * z))
        ux, uy, uz = r * np.cos(phi), r * np.sin(phi), z
    p = [(pscale * v).astype(T) for v in (ux, uy, uz)]
    return x[0], x[1], x[2], p[0], p[1], p[2]


def injection_rng(cfg):
    """Where init_particles draws: "host" (numpy PCG64, the default) or
    "device" (Philox4x64-10 counters per particle, pt_inject)."""
    m = _norm_mode(cfg.get("injection_rng", "host"))
    if m not in ("host", "device"):
        raise PipelineError("injection_rng must be host or device")
    return m


def inject_params(cfg, axes, gamma0, v0):
    """pt_inject_params for init_particles (phase_space_gpu_runner.jl:196-245)
    with the reference's checks on the settings (:210-219), in FP64."""
    from . import _lib

    ip = _lib.InjectParams()
    ip.seed = int(cfg["seed"]) & 0xFFFFFFFFFFFFFFFF
    fixed = injection_position_mode(cfg) == "fixed"
    ip.fixed_position = int(fixed)
    if fixed:
        for c, v in enumerate(fixed_injection_position(cfg, axes, np.float64)):
            ip.position_m[c] = float(v)
    fixed_mu = injection_mode(cfg) == "fixed_mu"
    ip.fixed_mu = int(fixed_mu)
    mu0 = float(cfg.get("injection_mu0", 0.0))
    if fixed_mu and not (-1.0 <= mu0 <= 1.0):
        raise PipelineError("injection_mu0 must be in [-1, 1]")
    ip.mu0 = mu0
    ip.p_magnitude = float(gamma0) * M_P * float(v0)
    return ip


def init_particles_device(cfg, ctx, field, axes, gamma0, v0, first_index=0, n=None):
    """init_particles (phase_space_gpu_runner.jl:196-245) drawn on the device
    by pt_inject for global particles first_index .. first_index+n-1 (n
    defaults to cfg["n_particles"]).  Fixed-mu injection interpolates B from
    the values the field stores.  Returns x1, x2, x3, p1, p2, p3 (FP64)."""
    import ctypes

    from . import _lib

    n = int(cfg["n_particles"]) if n is None else int(n)
    ip = inject_params(cfg, axes, gamma0, v0)
    out = np.empty((6, n), dtype=np.float64)
    _lib.check(_lib.load().pt_inject(ctx.h, field.h, int(first_index), n, ctypes.byref(ip), 0,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return tuple(out[c] for c in range(6))
