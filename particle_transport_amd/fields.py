"""Field inputs: the reference's loader and grid, and a synthetic turbulence generator.

``load_static_fields`` / ``build_uniform_coords`` / ``estimate_min_dx`` restate
phase_space_gpu_runner.jl:95-134.  The reference reads HDF5 datasets named by
``B_paths``/``v_paths``; the loader reads them from an HDF5 file through
libpt_h5 (h5py is not installed), from a ``.npz`` archive with the same names,
or takes arrays directly.

``generate_turbulence`` writes the synthetic inputs the benchmark configs
describe (the reference ships no generator): B = B0 z + dB and u = du, each
fluctuation an isotropic Gaussian random field whose shell spectrum is
|dB(k)|^2 ~ k^(-5/3) (3-D power ~ k^(-11/3)) for integer wavenumbers
|k| L / 2pi in [kmin, kmax], made divergence-free by dB(k) <- dB - k (k.dB)/k^2
and rescaled to the requested vector rms.  Fields are generated periodic on
(n-1)^3 points and the first plane is repeated as the last node plane, so the
node grid range(0, L, length=n) is seamless across the periodic wrap.
Values are in the reference's input units: B in microgauss, u in km/s.

``mode_decompose=True`` builds the mode-decomposed set instead (Cho &
Lazarian 2002, 2003): one seeded Gaussian vector field V with the same k^-5/3
spectrum (not made solenoidal) is projected, per k-mode, onto the
eigen-directions of the linear ideal-MHD wave operator with B0 along z and
c_s^2 / V_A^2 = alpha = (5/3) beta / 2: the Alfven direction xi_A = k x z / |k
x z| and the fast and slow directions xi_f, xi_s in the plane of k and z
(mhd_eigendirections).  Each projection V_m = (V.xi_m) xi_m is that mode's
displacement; its flow is u_m ~ V_m and its magnetic fluctuation the
displacement's field-line signature dB_m ~ i k x (V_m x z) / |k|, which is
perpendicular to k, so every magnetic part is divergence-free.  The parts sum
to the total (V and i k x (V x z) / |k|), both fluctuations are rescaled by one
factor each to the requested rms, and each mode field is B0 z + its own parts.
The total of a mode-decomposed set is therefore not the plain recipe's total
for the same seed.
"""
from __future__ import annotations

import math

import numpy as np

from .time_units import PipelineError

UGAUSS_TO_T = 1e-10
KMPS_TO_MPS = 1e3
PC_TO_M = 3.085677581e16

B_PATHS = ("i_mag_field", "j_mag_field", "k_mag_field")
V_PATHS = ("i_velocity", "j_velocity", "k_velocity")


def box_length_m(box_length_pc: float) -> float:
    """L = box_length_pc * PC_TO_M in Float64 (phase_space_gpu_runner.jl:126)."""
    return float(box_length_pc) * PC_TO_M


def build_uniform_coords(box_length_pc: float, nx: int, ny: int, nz: int):
    """Grid extents of collect(range(0, L, length=n)) (phase_space_gpu_runner.jl:125-130):
    xmin = 0, xmax = L, and x[2]-x[1] = L/(n-1) rounded once."""
    L = box_length_m(box_length_pc)
    return [dict(min=0.0, max=L, d=L / (n - 1), n=n) for n in (nx, ny, nz)]


def estimate_min_dx(axes) -> float:
    """phase_space_gpu_runner.jl:132-134 (uniform grid: the node spacing)."""
    return min(a["d"] for a in axes)


def load_static_fields(cfg, dtype=np.float64, arrays=None):
    """phase_space_gpu_runner.jl:95-123.  Returns (Bx, By, Bz, vx, vy, vz) as
    Fortran-ordered (nx,ny,nz) arrays in SI units (tesla, m/s).

    ``arrays`` may hold the six raw datasets (microgauss, velocity units);
    otherwise ``cfg["file"]`` names an HDF5 file (``.h5``/``.hdf5``) or a .npz
    archive with the B_paths/v_paths datasets, each a Julia (nx,ny,nz) array,
    i.e. the C-order (nz,ny,nx) dataspace HDF5 stores.  ``field_subset``
    (phase_space_gpu_runner.jl:100-104) keeps the first (nx,ny,nz) nodes; an
    HDF5 file is read as that box only."""
    subset = cfg.get("field_subset")
    bp = tuple(cfg.get("B_paths", B_PATHS))
    vp = tuple(cfg.get("v_paths", V_PATHS))
    if arrays is None and str(cfg["file"]).lower().endswith((".h5", ".hdf5")):
        from . import h5
        raw = []
        with h5.H5File(str(cfg["file"]), "r") as f:
            for k in (*bp, *vp):
                ds = f.dataset(k)
                if ds.ndim != 3:
                    raise ValueError(f"{k}: expected a 3-D field dataset, got shape {ds.shape}")
                count = tuple(ds.shape)
                if subset:
                    nx, ny, nz = _check_subset(subset, (count[2], count[1], count[0]), k)
                    count = (nz, ny, nx)
                raw.append(ds.read_box((0, 0, 0), count).transpose(2, 1, 0))
    elif arrays is None:
        with np.load(cfg["file"], allow_pickle=False) as z:
            raw = [np.asarray(z[k]).transpose(2, 1, 0) for k in (*bp, *vp)]
    else:
        raw = list(arrays)
    out = []
    for i, a in enumerate(raw):
        a = np.asarray(a)
        if subset:
            nx, ny, nz = _check_subset(subset, a.shape, (*bp, *vp)[i])
            a = a[:nx, :ny, :nz]
        a = np.asfortranarray(a, dtype=dtype).copy(order="F")
        scale = UGAUSS_TO_T if i < 3 else float(cfg.get("velocity_unit_in_m_per_s", KMPS_TO_MPS))
        a *= dtype(scale)
        out.append(a)
    return tuple(out)


def _check_subset(subset, shape, name):
    """field_subset (nx,ny,nz) against a dataset's Julia shape: the reference
    slices data[1:nx,1:ny,1:nz] (phase_space_gpu_runner.jl:100-104), which
    raises a BoundsError when the subset is larger than the dataset."""
    sub = tuple(int(v) for v in subset)
    if len(sub) != 3 or min(sub) < 1:
        raise PipelineError(f"field_subset must be three positive sizes, got {tuple(subset)}")
    if any(a > b for a, b in zip(sub, shape)):
        raise PipelineError(f"field_subset {sub} exceeds the shape {tuple(shape)} of dataset {name}")
    return sub


def _wavenumbers(M: int, xp, device=None):
    """Integer wavenumber grids for an rfftn over a C-order (M,M,M) array
    indexed [kz, ky, kx] (x fastest)."""
    if xp is np:
        k = np.fft.fftfreq(M, d=1.0 / M)
        kr = np.fft.rfftfreq(M, d=1.0 / M)
        KZ, KY, KX = np.meshgrid(k, k, kr, indexing="ij")
        return KX, KY, KZ
    import torch
    k = torch.fft.fftfreq(M, d=1.0 / M, device=device, dtype=torch.float64)
    kr = torch.fft.rfftfreq(M, d=1.0 / M, device=device, dtype=torch.float64)
    KZ, KY, KX = torch.meshgrid(k, k, kr, indexing="ij")
    return KX, KY, KZ


def _solenoidal_spectrum(M, kmin, kmax, rng, xp, device, solenoidal=True):
    """Three rfft components of a divergence-free (or, solenoidal=False,
    unprojected) Gaussian field with 3-D power ~ k^(-11/3) on the shell band
    [kmin, kmax]."""
    if xp is np:
        noise = rng.standard_normal((3, M, M, M))
        F = np.fft.rfftn(noise, axes=(1, 2, 3))
    else:
        import torch
        noise = torch.randn((3, M, M, M), generator=rng, device=device, dtype=torch.float32)
        F = torch.fft.rfftn(noise.double() if device == "cpu" else noise, dim=(1, 2, 3))
        del noise
    KX, KY, KZ = _wavenumbers(M, xp, device)
    k2 = KX * KX + KY * KY + KZ * KZ
    kmag = xp.sqrt(k2)
    band = (kmag >= kmin) & (kmag <= kmax)
    safe = xp.where(k2 > 0, k2, xp.ones_like(k2))
    amp = xp.where(band, safe ** (-11.0 / 12.0), xp.zeros_like(k2))
    F = F * amp
    if solenoidal:
        kdotF = (KX * F[0] + KY * F[1] + KZ * F[2]) / safe
        F = xp.stack([F[0] - KX * kdotF, F[1] - KY * kdotF, F[2] - KZ * kdotF])
    return F, (KX, KY, KZ, k2)


def mhd_eigendirections(KX, KY, KZ, alpha, xp=np):
    """Unit displacement directions of the three linear ideal-MHD waves for
    wavevectors (KX, KY, KZ) with B0 along z and c_s^2 / V_A^2 = alpha: the
    Alfven direction xi_A = k x z / |k x z| and the fast and slow directions
    in the plane of k and z, from Cho & Lazarian (2003, MNRAS 345, 325, eq. 3-4):
    xi_f ~ (1 + a + sqrt(D)) k_perp e_perp + (-1 + a + sqrt(D)) k_par z and
    xi_s ~ (1 + a - sqrt(D)) k_perp e_perp + (-1 + a - sqrt(D)) k_par z,
    D = (1 + a)^2 - 4 a cos^2(theta), with k_perp, k_par the components of k/|k|.
    Where one of the two vanishes (k along z) its direction is taken as the
    in-plane perpendicular of the other (where both do, k along z at alpha =
    1, fast = e_perp and slow = z); k along z also takes xi_A = x and e_perp
    = y.  Returns three (x, y, z) component tuples; the sign of each
    direction is arbitrary (projections do not depend on it)."""
    k2 = KX * KX + KY * KY + KZ * KZ
    one = xp.ones_like(k2)
    zero = xp.zeros_like(k2)
    kmag = xp.sqrt(xp.where(k2 > 0, k2, one))
    kperp_abs = xp.sqrt(KX * KX + KY * KY)
    has_perp = kperp_abs > 0
    kp_safe = xp.where(has_perp, kperp_abs, one)
    eA = (xp.where(has_perp, KY / kp_safe, one), xp.where(has_perp, -KX / kp_safe, zero), zero)
    ep = (xp.where(has_perp, KX / kp_safe, zero), xp.where(has_perp, KY / kp_safe, one), zero)
    kperp = kperp_abs / kmag
    kpar = KZ / kmag
    sqD = xp.sqrt(xp.clip((1 + alpha) ** 2 - 4 * alpha * kpar * kpar, 0, None))
    # in-plane coordinates (along e_perp, along z) of the two directions
    fp, fz = (1 + alpha + sqD) * kperp, (-1 + alpha + sqD) * kpar
    sp, sz = (1 + alpha - sqD) * kperp, (-1 + alpha - sqD) * kpar
    fn = xp.sqrt(fp * fp + fz * fz)
    sn = xp.sqrt(sp * sp + sz * sz)
    use_f = fn >= sn
    # normalise the larger one, take the other as its in-plane perpendicular
    big_p = xp.where(use_f, fp, sp)
    big_z = xp.where(use_f, fz, sz)
    bn = xp.sqrt(big_p * big_p + big_z * big_z)
    # both vanish only for k along z at alpha = 1, where the two waves are
    # degenerate: fast transverse (e_perp), slow along z
    degenerate = bn <= 0
    big_p = xp.where(degenerate, one, big_p / xp.where(degenerate, one, bn))
    big_z = xp.where(degenerate, zero, big_z / xp.where(degenerate, one, bn))
    use_f = use_f | degenerate
    fp_u = xp.where(use_f, big_p, big_z)
    fz_u = xp.where(use_f, big_z, -big_p)
    sp_u = xp.where(use_f, -big_z, big_p)
    sz_u = xp.where(use_f, big_p, big_z)
    xi_f = (fp_u * ep[0], fp_u * ep[1], fz_u)
    xi_s = (sp_u * ep[0], sp_u * ep[1], sz_u)
    return eA, xi_f, xi_s


def _mode_parts(V, K, xp, alpha):
    """Per mode (alfven, fast, slow): the flow part V_m = (V.xi_m) xi_m and the
    magnetic part i k x (V_m x z) / |k| = i (V_m k_z - z (k.V_m)) / |k|."""
    KX, KY, KZ, k2 = K
    kmag = xp.sqrt(xp.where(k2 > 0, k2, xp.ones_like(k2)))
    khx, khy, khz = KX / kmag, KY / kmag, KZ / kmag
    dirs = dict(zip(("alfven", "fast", "slow"), mhd_eigendirections(KX, KY, KZ, alpha, xp)))
    out = {}
    for name, e in dirs.items():
        a = V[0] * e[0] + V[1] * e[1] + V[2] * e[2]
        flow = xp.stack([a * e[0], a * e[1], a * e[2]])
        kdot = khx * flow[0] + khy * flow[1] + khz * flow[2]
        mag = xp.stack([flow[0] * khz, flow[1] * khz, flow[2] * khz - kdot]) * 1j
        out[name] = (mag, flow)
    return out


def _to_real(F, M, xp):
    if xp is np:
        return np.fft.irfftn(F, s=(M, M, M), axes=(1, 2, 3))
    import torch
    return torch.fft.irfftn(F, s=(M, M, M), dim=(1, 2, 3))


def _with_seam(r, xp):
    """(3,M,M,M) periodic -> (3,M+1,M+1,M+1) node grid with the first plane repeated."""
    if xp is np:
        return np.pad(r, ((0, 0), (0, 1), (0, 1), (0, 1)), mode="wrap")
    import torch
    r = torch.cat([r, r[:, :1]], dim=1)
    r = torch.cat([r, r[:, :, :1]], dim=2)
    return torch.cat([r, r[:, :, :, :1]], dim=3)


def generate_turbulence(n_nodes: int, *, B0_uG: float = 5.0, dB_over_B0: float = 0.5, u_rms_kms: float = 20.0,
                        kmin: float = 1.0, kmax: float = 16.0, seed: int = 1234, device=None,
                        dtype="float64", mode_decompose: bool = False, mode_beta: float = 1.0):
    """Synthetic B (microgauss) and u (km/s) on an n_nodes^3 node grid.

    Returns a dict with key "total" (and "alfven"/"fast"/"slow" when
    mode_decompose) mapping to a 6-tuple (Bx, By, Bz, ux, uy, uz), each shaped
    (nx, ny, nz) with x fastest in memory (the Julia layout).  device=None
    uses numpy; a torch device string ("cuda", "cpu") uses torch.fft there.
    mode_beta: plasma beta of the mode decomposition (c_s^2/V_A^2 = 5/6 beta).
    """
    M = int(n_nodes) - 1
    if M < 2:
        raise ValueError("need at least 3 nodes per axis")
    if device is None:
        xp, rng = np, np.random.default_rng(seed)
    else:
        import torch
        xp = torch
        rng = torch.Generator(device=device)
        rng.manual_seed(int(seed))
    def realise(Fc):
        return _to_real(Fc, M, xp)

    if mode_decompose:
        V, K = _solenoidal_spectrum(M, kmin, kmax, rng, xp, device, solenoidal=False)
        parts = _mode_parts(V, K, xp, 5.0 / 6.0 * float(mode_beta))
        del V
        real_parts = {name: (realise(mag), realise(flow)) for name, (mag, flow) in parts.items()}
        del parts
        rB = sum(b for b, _ in real_parts.values())
        rU = sum(u for _, u in real_parts.values())
    else:
        FB, K = _solenoidal_spectrum(M, kmin, kmax, rng, xp, device)
        FU, _ = _solenoidal_spectrum(M, kmin, kmax, rng, xp, device)
        rB = realise(FB)
        rU = realise(FU)
        del FB, FU
    sB = float(dB_over_B0 * B0_uG) / float(xp.sqrt((rB * rB).sum(0).mean()))
    sU = float(u_rms_kms) / float(xp.sqrt((rU * rU).sum(0).mean()))

    def pack(rb, ru):
        b = _with_seam(rb * sB, xp)
        u = _with_seam(ru * sU, xp)
        comps = [b[0], b[1], b[2] + B0_uG, u[0], u[1], u[2]]
        out = []
        for c in comps:
            if xp is np:
                c = np.ascontiguousarray(c, dtype=np.dtype(dtype))
                out.append(c.transpose(2, 1, 0))   # (nx,ny,nz) view, x fastest
            else:
                c = c.to(getattr(xp, dtype)).contiguous()
                out.append(c.permute(2, 1, 0))
        return tuple(out)

    result = {"total": pack(rB, rU)}
    if mode_decompose:
        for name in ("alfven", "fast", "slow"):
            result[name] = pack(*real_parts[name])
    return result


def to_si(fields, velocity_unit_in_m_per_s=KMPS_TO_MPS):
    """Scale (B in microgauss, u in velocity units) to (tesla, m/s), as
    load_static_fields does (phase_space_gpu_runner.jl:114-119)."""
    out = []
    for i, a in enumerate(fields):
        out.append(a * (UGAUSS_TO_T if i < 3 else velocity_unit_in_m_per_s))
    return tuple(out)


def _generate_si(n_nodes, *, seed, dB_over_B0, kmin, kmax, B0_uG, u_rms_kms, mode_decompose, device):
    """Yield (field index, component index, (nz,ny,nx) float32 numpy array) in SI units."""
    if device == "auto":
        try:
            import torch
            device = "cuda" if torch.cuda.is_available() else None
        except ImportError:
            device = None
    gen = generate_turbulence(n_nodes, B0_uG=B0_uG, dB_over_B0=dB_over_B0, u_rms_kms=u_rms_kms, kmin=kmin,
                              kmax=kmax, seed=seed, device=device, dtype="float32", mode_decompose=mode_decompose)
    names = ("total", "alfven", "fast", "slow") if mode_decompose else ("total",)
    for f, name in enumerate(names):
        for c, a in enumerate(gen[name]):
            scale = UGAUSS_TO_T if c < 3 else KMPS_TO_MPS
            if device is None:
                yield f, c, (np.asarray(a).transpose(2, 1, 0) * scale).astype(np.float32)
            else:
                yield f, c, (a.permute(2, 1, 0) * scale).float().cpu().numpy()


def generate_to_stream(stream, n_nodes, **kw):
    """Write the synthetic fields in SI units (tesla, m/s) to a binary stream:
    one header line "n_fields n_nodes\n", then float32 arrays (nz,ny,nx) for
    field 0 components 0..5, field 1 ..., x fastest.  Field 0 is the total
    field; with mode_decompose fields 1-3 are Alfven/fast/slow.

    Meant for a child process (``python -m particle_transport_amd.fields``)
    when torch runs the FFTs: torch carries its own HIP runtime, which cannot
    share a process with libpt_hip.so's."""
    nf = 4 if kw.get("mode_decompose") else 1
    stream.write(f"{nf} {n_nodes}\n".encode())
    for _, _, a in _generate_si(n_nodes, **kw):
        stream.write(np.ascontiguousarray(a).data)
    stream.flush()


def read_stream(stream):
    """Read generate_to_stream's output: a list (one entry per field) of
    6-tuples of (nx,ny,nz) Fortran-ordered float32 arrays."""
    nf, M = (int(v) for v in stream.readline().split())
    out = np.empty((nf, 6, M, M, M), dtype=np.float32)
    view = memoryview(out.reshape(-1)).cast("B")
    got = 0
    while got < len(view):
        k = stream.readinto(view[got:got + (1 << 26)])
        if not k:
            raise EOFError("field stream ended early")
        got += k
    return [tuple(out[f, c].transpose(2, 1, 0) for c in range(6)) for f in range(nf)]


if __name__ == "__main__":
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="write synthetic turbulence fields (SI, float32) to stdout")
    ap.add_argument("--nodes", type=int, required=True)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--dB", type=float, default=1.0)
    ap.add_argument("--kmin", type=float, default=1.0)
    ap.add_argument("--kmax", type=float, default=64.0)
    ap.add_argument("--modes", action="store_true")
    ap.add_argument("--cpu", action="store_true")
    a = ap.parse_args()
    generate_to_stream(sys.stdout.buffer, a.nodes, seed=a.seed, dB_over_B0=a.dB, kmin=a.kmin, kmax=a.kmax,
                       B0_uG=5.0, u_rms_kms=20.0, mode_decompose=a.modes, device=None if a.cpu else "auto")
