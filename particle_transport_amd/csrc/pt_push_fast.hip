// FP32 split-precision pushers for gfx950.
//
//   PT_PUSH_PLAIN_F32  the reference's structure in FP32: one launch per step,
//                      state read from and written to HBM each step, the six
//                      planar field arrays gathered at all 8 corners every
//                      step with the reference's lerp form
//                      (trilinear_sample, phase_space_gpu_runner.jl:247-276).
//   PT_PUSH_FAST       the production pusher: K steps per launch, one particle
//                      per lane, state in registers, and the current cell's
//                      trilinear coefficients (6 components x 8) cached in
//                      registers and reloaded only when the particle changes
//                      cell.  Fields are 16-byte BF16/FP16 node records (or
//                      32-byte FP32 records) in 8^3 bricks, so a reload is 8
//                      vector loads.
//
// Both use split position (int32 cell + FP32 offset in [0,1) cell units, so
// periodic wrap is an exact integer operation) and the momentum in units of
// m c.  The Boris step is the reference's (boris_push, :282-323) with the
// physical factors folded into the fields:
//     B~ = (q dt / 2m) B,  u~ = u / c,  E~ = -u~ x B~ = (q dt / 2mc) E,
//     p-  = P + E~,  t = B~ / gamma-,  s = 2t/(1+t^2),
//     p'  = p- + p- x t,  p+ = p- + p' x s,  P_new = p+ + E~,
//     offset += P_new (c dt / dx) / gamma_new.
#include <algorithm>

#include "pt_internal.hpp"

namespace pt {
namespace {

struct FieldView {
    const void *data;
    int64_t n[3];
    int64_t nb[3];
    float scale[6];
    float offset[6];
};

struct SplitGeom {
    int32_t nc[3];   // cells per axis (nodes - 1)
    float k[3];      // c dt / dx per axis
    double d[3];     // node spacing (m), for phase-space output
};

struct Phys {
    float fac[6];    // B: q dt / 2m ; u: 1/c
    float qsign;     // sign of the charge (mu is defined against B, not B~)
    float mr;        // M_P / m : state momentum (units M_P c) -> units m c
    double pout;     // m c : momentum in units m c -> kg m/s
};

// ------------------------------------------------------- shared math -------

__device__ __forceinline__ void boris_split(float &Ux, float &Uy, float &Uz, float bx, float by,
                                            float bz, float ux, float uy, float uz, float &inv_gn) {
    // E~ = -(u~ x B~)
    const float ex = -(uy * bz - uz * by);
    const float ey = -(uz * bx - ux * bz);
    const float ez = -(ux * by - uy * bx);
    const float pmx = Ux + ex, pmy = Uy + ey, pmz = Uz + ez;
    const float gm = sqrtf(1.0f + (pmx * pmx + pmy * pmy + pmz * pmz));
    const float r = 1.0f / gm;
    const float tx = bx * r, ty = by * r, tz = bz * r;
    const float t2 = tx * tx + ty * ty + tz * tz;
    const float f = 2.0f / (1.0f + t2);
    const float sx = tx * f, sy = ty * f, sz = tz * f;
    const float ppx = pmx + (pmy * tz - pmz * ty);
    const float ppy = pmy + (pmz * tx - pmx * tz);
    const float ppz = pmz + (pmx * ty - pmy * tx);
    Ux = pmx + (ppy * sz - ppz * sy) + ex;
    Uy = pmy + (ppz * sx - ppx * sz) + ey;
    Uz = pmz + (ppx * sy - ppy * sx) + ez;
    inv_gn = 1.0f / sqrtf(1.0f + (Ux * Ux + Uy * Uy + Uz * Uz));
}

// Advance one axis of the split position by du cells; returns false if an
// open boundary kills the particle.
__device__ __forceinline__ void move_axis(int32_t &cell, float &off, float du, int32_t nc,
                                          bool periodic) {
    off += du;
    if (!(off >= 0.0f && off < 1.0f)) {
        const float fl = floorf(off);
        cell += static_cast<int32_t>(fl);
        off -= fl;
        if (off >= 1.0f) {  // -tiny + 1 rounds to 1: stay put at offset 0
            off -= 1.0f;
            cell += 1;
        }
        if (periodic && (cell < 0 || cell >= nc)) {
            cell %= nc;
            if (cell < 0) cell += nc;
        }
    }
}

// x > L or x < 0 kills (phase_space_gpu_runner.jl:385-387); x == L is cell nc, offset 0.
__device__ __forceinline__ bool outside(int32_t cell, float off, int32_t nc) {
    return cell < 0 || cell > nc || (cell == nc && off > 0.0f);
}

// Gather cell and fractional coordinate: x == L gathers the last cell at t = 1
// (the clamp of i0 to nx-1 in trilinear_sample).
__device__ __forceinline__ void gather_coord(int32_t cell, float off, int32_t nc, int32_t &g,
                                             float &t) {
    g = cell;
    t = off;
    if (g > nc - 1) {
        t += static_cast<float>(g - (nc - 1));
        g = nc - 1;
    }
    // A live particle always has 0 <= cell <= nc; the clamp keeps every field
    // access in bounds even for a state that is not (e.g. non-finite input).
    g = g < 0 ? 0 : g;
}

template <class T>
__device__ __forceinline__ void put(void *out, int64_t idx, double v) {
    static_cast<T *>(out)[idx] = static_cast<T>(v);
}

__device__ __forceinline__ float mu_f32(float px, float py, float pz, float bx, float by, float bz,
                                        float qsign) {
    const float pn = sqrtf(px * px + py * py + pz * pz);
    const float bn = sqrtf(bx * bx + by * by + bz * bz);
    if (!isfinite(pn) || !isfinite(bn) || pn == 0.0f || bn == 0.0f) return NAN;
    const float mu = qsign * (px * bx + py * by + pz * bz) / (pn * bn);
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// Save-step output of one particle: state before the push of step si, as
// the reference's *_step copies (:351-356) or mu (pipeline :1349-1374).
__device__ __forceinline__ void write_save(int out_kind, int out_f64, void *out_a, void *out_b,
                                           int64_t slot, int64_t n, int64_t p, bool alive,
                                           const int32_t cell[3], const float off[3],
                                           const float U[3], const float bt[3],
                                           const SplitGeom &g, const Phys &ph) {
    if (out_kind == PT_OUT_PHASE) {
        for (int c = 0; c < 3; ++c) {
            const double xv = alive ? (static_cast<double>(cell[c]) + static_cast<double>(off[c])) * g.d[c] : NAN;
            const double pv = alive ? static_cast<double>(U[c]) * ph.pout : NAN;
            const int64_t idx = (slot * 3 + c) * n + p;
            if (out_f64) {
                put<double>(out_a, idx, xv);
                put<double>(out_b, idx, pv);
            } else {
                put<float>(out_a, idx, xv);
                put<float>(out_b, idx, pv);
            }
        }
    } else {
        const float mu = alive ? mu_f32(U[0], U[1], U[2], bt[0], bt[1], bt[2], ph.qsign) : NAN;
        if (out_f64)
            put<double>(out_a, slot * n + p, mu);
        else
            put<float>(out_a, slot * n + p, mu);
    }
}

// Map hardware workgroup ids (dispatched round-robin over the 8 XCDs) so that
// each XCD processes one contiguous range of particles, keeping a spatially
// sorted ensemble's field records in that XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblk) {
    const int64_t xcd = b & 7, q = nblk >> 3, r = nblk & 7;
    return xcd * q + (xcd < r ? xcd : r) + (b >> 3);
}

// ----------------------------------------------------------- plain ---------

__global__ void __launch_bounds__(256)
k_push_plain(SplitState s, uint8_t *__restrict__ alive_arr, const float *__restrict__ F,
             int64_t nx, int64_t ny, int64_t nz, SplitGeom g, Phys ph, int64_t n, int64_t si,
             int64_t total_steps, int64_t stride, int64_t save_base, int out_kind, int out_f64,
             void *out_a, void *out_b, int periodic, int64_t first_step,
             unsigned long long *deaths) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const bool save = out_kind != PT_OUT_NONE && ((si - 1) % stride) == 0;
    bool alive = alive_arr[p] != 0;
    const bool push = alive && si < total_steps;
    if (!save && !push) return;
    int32_t cell[3];
    float off[3], U[3];
    for (int c = 0; c < 3; ++c) {
        cell[c] = s.cell[c][p];
        off[c] = s.off[c][p];
        U[c] = s.P[c][p] * ph.mr;
    }
    float v[6] = {0, 0, 0, 0, 0, 0};
    if (alive) {
        int32_t gi, gj, gk;
        float tx, ty, tz;
        gather_coord(cell[0], off[0], g.nc[0], gi, tx);
        gather_coord(cell[1], off[1], g.nc[1], gj, ty);
        gather_coord(cell[2], off[2], g.nc[2], gk, tz);
        const int64_t nn = nx * ny * nz;
        const int64_t base = gi + nx * (gj + ny * static_cast<int64_t>(gk));
        const int64_t sy = nx, sz = nx * ny;
        // All six components at all eight corners, every step, as the reference.
        for (int c = 0; c < 6; ++c) {
            const float *arr = F + c * nn;
            const float c000 = arr[base], c100 = arr[base + 1];
            const float c010 = arr[base + sy], c110 = arr[base + 1 + sy];
            const float c001 = arr[base + sz], c101 = arr[base + 1 + sz];
            const float c011 = arr[base + sy + sz], c111 = arr[base + 1 + sy + sz];
            const float c00 = c000 * (1 - tx) + c100 * tx;
            const float c10 = c010 * (1 - tx) + c110 * tx;
            const float c01 = c001 * (1 - tx) + c101 * tx;
            const float c11 = c011 * (1 - tx) + c111 * tx;
            const float c0 = c00 * (1 - ty) + c10 * ty;
            const float c1 = c01 * (1 - ty) + c11 * ty;
            v[c] = (c0 * (1 - tz) + c1 * tz) * ph.fac[c];
        }
    }
    if (save) {
        const int64_t slot = (si - 1) / stride - save_base;
        write_save(out_kind, out_f64, out_a, out_b, slot, n, p, alive, cell, off, U, v, g, ph);
    }
    if (!push) return;
    float inv_gn;
    boris_split(U[0], U[1], U[2], v[0], v[1], v[2], v[3], v[4], v[5], inv_gn);
    for (int c = 0; c < 3; ++c) move_axis(cell[c], off[c], U[c] * (g.k[c] * inv_gn), g.nc[c], periodic);
    if (!periodic && (outside(cell[0], off[0], g.nc[0]) || outside(cell[1], off[1], g.nc[1]) ||
                      outside(cell[2], off[2], g.nc[2]))) {
        alive = false;
        atomicAdd(&deaths[si - first_step], 1ull);
        alive_arr[p] = 0;
    }
    for (int c = 0; c < 3; ++c) {
        s.cell[c][p] = cell[c];
        s.off[c][p] = off[c];
        s.P[c][p] = U[c] / ph.mr;
    }
}

// ------------------------------------------------------------ fast ---------

// Corner values of the six components at the 8 nodes of cell (gi,gj,gk),
// corner index ci = dx + 2 dy + 4 dz, in storage units.
template <int STORE>
__device__ __forceinline__ void load_corners(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk,
                                             float s[6][8]) {
#pragma unroll
    for (int ci = 0; ci < 8; ++ci) {
        const int64_t i = gi + (ci & 1), j = gj + ((ci >> 1) & 1), k = gk + (ci >> 2);
        if constexpr (STORE == PT_STORE_F32_PLANAR) {
            const int64_t nn = fv.n[0] * fv.n[1] * fv.n[2];
            const int64_t idx = i + fv.n[0] * (j + fv.n[1] * k);
            const float *F = static_cast<const float *>(fv.data);
#pragma unroll
            for (int c = 0; c < 6; ++c) s[c][ci] = F[c * nn + idx];
        } else {
            const int64_t rec = brick_node_index(fv.nb, i, j, k);
            if constexpr (STORE == PT_STORE_F32_BRICK) {
                const float4 *R = static_cast<const float4 *>(fv.data) + rec * 2;
                const float4 a = R[0];
                const float4 b = R[1];
                s[0][ci] = a.x;
                s[1][ci] = a.y;
                s[2][ci] = a.z;
                s[3][ci] = a.w;
                s[4][ci] = b.x;
                s[5][ci] = b.y;
            } else {
                const uint4 w = static_cast<const uint4 *>(fv.data)[rec];
                if constexpr (STORE == PT_STORE_BF16_BRICK) {
                    s[0][ci] = __uint_as_float(w.x << 16);
                    s[1][ci] = __uint_as_float(w.x & 0xffff0000u);
                    s[2][ci] = __uint_as_float(w.y << 16);
                    s[3][ci] = __uint_as_float(w.y & 0xffff0000u);
                    s[4][ci] = __uint_as_float(w.z << 16);
                    s[5][ci] = __uint_as_float(w.z & 0xffff0000u);
                } else {
                    s[0][ci] = f16_bits_to_float(static_cast<uint16_t>(w.x & 0xffffu));
                    s[1][ci] = f16_bits_to_float(static_cast<uint16_t>(w.x >> 16));
                    s[2][ci] = f16_bits_to_float(static_cast<uint16_t>(w.y & 0xffffu));
                    s[3][ci] = f16_bits_to_float(static_cast<uint16_t>(w.y >> 16));
                    s[4][ci] = f16_bits_to_float(static_cast<uint16_t>(w.z & 0xffffu));
                    s[5][ci] = f16_bits_to_float(static_cast<uint16_t>(w.z >> 16));
                }
            }
        }
    }
}

// Trilinear polynomial coefficients of one component over a cell:
// f(x,y,z) = a0 + a1 x + a2 y + a3 z + a4 xy + a5 xz + a6 yz + a7 xyz,
// decoded (stored*scale + offset) and multiplied by the physical factor.
__device__ __forceinline__ void coeffs(const float s[8], float scale, float offset, float a[8]) {
    const float dx00 = s[1] - s[0], dx10 = s[3] - s[2], dx01 = s[5] - s[4], dx11 = s[7] - s[6];
    const float dy0 = s[2] - s[0], dy1 = s[6] - s[4];
    a[0] = s[0] * scale + offset;
    a[1] = dx00 * scale;
    a[2] = dy0 * scale;
    a[3] = (s[4] - s[0]) * scale;
    a[4] = (dx10 - dx00) * scale;
    a[5] = (dx01 - dx00) * scale;
    a[6] = (dy1 - dy0) * scale;
    a[7] = ((dx11 - dx01) - (dx10 - dx00)) * scale;
}

__device__ __forceinline__ float eval_poly(const float a[8], float x, float y, float z) {
    const float z0 = fmaf(a[3], z, a[0]);
    const float z1 = fmaf(a[5], z, a[1]);
    const float z2 = fmaf(a[6], z, a[2]);
    const float z3 = fmaf(a[7], z, a[4]);
    return fmaf(fmaf(z3, x, z2), y, fmaf(z1, x, z0));
}

template <int STORE>
__global__ void __launch_bounds__(256)
k_push_fast(SplitState s, uint8_t *__restrict__ alive_arr, FieldView fv, SplitGeom g, Phys ph,
            int64_t n, int64_t si0, int K, int64_t total_steps, int64_t stride, int64_t save_base,
            int out_kind, int out_f64, void *out_a, void *out_b, int periodic, int64_t first_step,
            unsigned long long *deaths) {
    const int64_t nblk = gridDim.x;
    const int64_t p = xcd_block(blockIdx.x, nblk) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    bool alive = alive_arr[p] != 0;
    int32_t cell[3];
    float off[3], U[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cell[c] = s.cell[c][p];
        off[c] = s.off[c][p];
        U[c] = s.P[c][p] * ph.mr;
    }
    int32_t cg0 = -1, cg1 = -1, cg2 = -1;
    float a[6][8];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int q = 0; q < 8; ++q) a[c][q] = 0.0f;

    for (int k = 0; k < K; ++k) {
        const int64_t si = si0 + k;
        const bool save = out_kind != PT_OUT_NONE && ((si - 1) % stride) == 0;
        const bool push = alive && si < total_steps;
        if (!save && !push) continue;
        float v[6] = {0, 0, 0, 0, 0, 0};
        if (alive) {
            int32_t gi, gj, gk;
            float tx, ty, tz;
            gather_coord(cell[0], off[0], g.nc[0], gi, tx);
            gather_coord(cell[1], off[1], g.nc[1], gj, ty);
            gather_coord(cell[2], off[2], g.nc[2], gk, tz);
            if (gi != cg0 || gj != cg1 || gk != cg2) {
                float sv[6][8];
                load_corners<STORE>(fv, gi, gj, gk, sv);
#pragma unroll
                for (int c = 0; c < 6; ++c) coeffs(sv[c], fv.scale[c] * ph.fac[c], fv.offset[c] * ph.fac[c], a[c]);
                cg0 = gi;
                cg1 = gj;
                cg2 = gk;
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) v[c] = eval_poly(a[c], tx, ty, tz);
        }
        if (save) {
            const int64_t slot = (si - 1) / stride - save_base;
            write_save(out_kind, out_f64, out_a, out_b, slot, n, p, alive, cell, off, U, v, g, ph);
        }
        if (!push) continue;
        float inv_gn;
        boris_split(U[0], U[1], U[2], v[0], v[1], v[2], v[3], v[4], v[5], inv_gn);
#pragma unroll
        for (int c = 0; c < 3; ++c) move_axis(cell[c], off[c], U[c] * (g.k[c] * inv_gn), g.nc[c], periodic);
        if (!periodic && (outside(cell[0], off[0], g.nc[0]) || outside(cell[1], off[1], g.nc[1]) ||
                          outside(cell[2], off[2], g.nc[2]))) {
            alive = false;
            atomicAdd(&deaths[si - first_step], 1ull);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.cell[c][p] = cell[c];
        s.off[c][p] = off[c];
        s.P[c][p] = U[c] / ph.mr;
    }
    alive_arr[p] = alive ? 1 : 0;
}

void make_geom_phys(const pt_field *f, const PushArgs &a, SplitGeom &g, Phys &ph) {
    for (int c = 0; c < 3; ++c) {
        g.nc[c] = static_cast<int32_t>(f->g.n[c] - 1);
        g.k[c] = static_cast<float>(kC * a.dt / f->g.d[c]);
        g.d[c] = f->g.d[c];
    }
    const double bfac = a.q * a.dt / (2.0 * a.m);
    for (int c = 0; c < 3; ++c) {
        ph.fac[c] = static_cast<float>(bfac);
        ph.fac[3 + c] = static_cast<float>(1.0 / kC);
    }
    ph.qsign = a.q > 0 ? 1.0f : -1.0f;
    ph.mr = static_cast<float>(kMp / a.m);
    ph.pout = a.m * kC;
}

}  // namespace

int64_t launch_push_plain_f32(pt_ens *e, const pt_field *f, const PushArgs &a) {
    SplitGeom g;
    Phys ph;
    make_geom_phys(f, a, g, ph);
    const unsigned blocks = static_cast<unsigned>((e->n + 255) / 256);
    int64_t launches = 0;
    for (int64_t k = 0; k < a.n_steps; ++k) {
        k_push_plain<<<blocks, 256, 0, e->ctx->stream>>>(
            split_state(e), e->alive, static_cast<const float *>(f->data), f->g.n[0], f->g.n[1], f->g.n[2],
            g, ph, e->n, a.first_step + k, a.total_steps, a.stride, a.save_base, a.out_kind, a.out_f64,
            a.out_a, a.out_b, e->periodic, a.first_step, reinterpret_cast<unsigned long long *>(e->deaths));
        PT_HIP_CHECK(hipGetLastError());
        ++launches;
    }
    return launches;
}

int64_t launch_push_fast(pt_ens *e, const pt_field *f, const PushArgs &a) {
    SplitGeom g;
    Phys ph;
    make_geom_phys(f, a, g, ph);
    FieldView fv;
    fv.data = f->data;
    for (int c = 0; c < 3; ++c) {
        fv.n[c] = f->g.n[c];
        fv.nb[c] = f->g.nb[c];
    }
    for (int c = 0; c < 6; ++c) {
        fv.scale[c] = f->sc.scale[c];
        fv.offset[c] = f->sc.offset[c];
    }
    const int K = a.steps_per_launch > 0 ? a.steps_per_launch : 128;
    const unsigned blocks = static_cast<unsigned>((e->n + 255) / 256);
    int64_t launches = 0;
    for (int64_t done = 0; done < a.n_steps; done += K) {
        const int k = static_cast<int>(std::min<int64_t>(K, a.n_steps - done));
        auto *dl = reinterpret_cast<unsigned long long *>(e->deaths);
        const int64_t si0 = a.first_step + done;
        switch (f->storage) {
            case PT_STORE_F32_PLANAR:
                k_push_fast<PT_STORE_F32_PLANAR><<<blocks, 256, 0, e->ctx->stream>>>(
                    split_state(e), e->alive, fv, g, ph, e->n, si0, k, a.total_steps, a.stride, a.save_base,
                    a.out_kind, a.out_f64, a.out_a, a.out_b, e->periodic, a.first_step, dl);
                break;
            case PT_STORE_F32_BRICK:
                k_push_fast<PT_STORE_F32_BRICK><<<blocks, 256, 0, e->ctx->stream>>>(
                    split_state(e), e->alive, fv, g, ph, e->n, si0, k, a.total_steps, a.stride, a.save_base,
                    a.out_kind, a.out_f64, a.out_a, a.out_b, e->periodic, a.first_step, dl);
                break;
            case PT_STORE_BF16_BRICK:
                k_push_fast<PT_STORE_BF16_BRICK><<<blocks, 256, 0, e->ctx->stream>>>(
                    split_state(e), e->alive, fv, g, ph, e->n, si0, k, a.total_steps, a.stride, a.save_base,
                    a.out_kind, a.out_f64, a.out_a, a.out_b, e->periodic, a.first_step, dl);
                break;
            case PT_STORE_F16_BRICK:
                k_push_fast<PT_STORE_F16_BRICK><<<blocks, 256, 0, e->ctx->stream>>>(
                    split_state(e), e->alive, fv, g, ph, e->n, si0, k, a.total_steps, a.stride, a.save_base,
                    a.out_kind, a.out_f64, a.out_a, a.out_b, e->periodic, a.first_step, dl);
                break;
            default:
                throw Error{PT_ERR_INVALID, "PT_PUSH_FAST: unsupported field storage"};
        }
        PT_HIP_CHECK(hipGetLastError());
        ++launches;
    }
    return launches;
}

}  // namespace pt
