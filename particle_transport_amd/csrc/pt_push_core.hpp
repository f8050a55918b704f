// Device core shared by the FP32 split-precision pushers (pt_push_fast.hip)
// and the multi-campaign serving kernel (pt_serve.hip): field views, the
// Boris step in folded units, split-position moves, corner loads and the
// per-cell trilinear polynomial.  See pt_push_fast.hip for the math.
#pragma once

#include "pt_internal.hpp"

// Multiply-adds are fused only inside one source expression (and where fmaf
// is written), never across statements by the backend: the same core then
// rounds identically in every kernel that inlines it (the single-campaign
// pusher and the serving kernel give bit-identical trajectories).  In effect
// for the rest of each translation unit that includes this header.
#pragma clang fp contract(on)

namespace pt {
namespace core {

struct FieldView {
    const void *data;
    int64_t n[3];
    int64_t nb[3];
    float scale[6];       // decoded value = stored * scale (+ the coarse part, 16-bit bricks)
    const float *coarse;  // 16-bit bricks: coarse-part coefficients [coarse cell][6][8]
    int32_t cn[3];        // coarse cells per axis
    const float *cells;   // cell table (kStoreCells kernels), else null
    int32_t cb[3];        // cell bricks per axis
};

__host__ inline FieldView make_field_view(const pt_field *f) {
    FieldView fv;
    fv.data = f->data;
    for (int c = 0; c < 3; ++c) {
        fv.n[c] = f->g.n[c];
        fv.nb[c] = f->g.nb[c];
        fv.cn[c] = f->cn[c];
    }
    for (int c = 0; c < 6; ++c) fv.scale[c] = f->sc.scale[c];
    fv.coarse = f->coarse;
    fv.cells = f->cells;
    for (int c = 0; c < 3; ++c) fv.cb[c] = f->cb[c];
    return fv;
}

template <int STORE>
constexpr bool kHierarchical = STORE == PT_STORE_BF16_BRICK || STORE == PT_STORE_F16_BRICK;

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

// One Boris push in folded units with the momentum magnitude carried in FP64.
// U (FP32 vector, units m c) and P2 = |U|^2 (FP64): the two half E-kicks
// are the only steps that change |p|.  With U- = U + e, U' its rotation
// (|U'| = |U-|) and U+ = U' + e, |U+|^2 - |U|^2 = 2 (U + U+).e exactly, so
// P2 is advanced by that FP32 dot product (about 1e-5 of |U|^2) in one FP64
// add, the rotation is applied to the FP32 vector, and the result is
// rescaled to |U| = sqrt(P2).  Without this the FP32 rotation's rounding
// drifts |p| by ~5e-8 per push, systematically.
//
// `renorm` (uniform across the wave) applies the rescale; between rescales
// |U| is off by at most the rotation drift of the pushes since the last one.
// The Lorentz factors come from the FP32 vector (1/gamma- from 1 + |U-|^2,
// 1/gamma+ from 1 + |U+|^2 before the rescale, both by the hardware
// reciprocal square root), so they depend only on the particle's state,
// never on how the steps are split into launches.  inv_gn returns 1/gamma+.
__device__ __forceinline__ void boris_split(float &Ux, float &Uy, float &Uz, double &P2, float bx, float by,
                                            float bz, float ux, float uy, float uz, float &inv_gn,
                                            bool renorm = true) {
    // E~ = -(u~ x B~)
    const float ex = -(uy * bz - uz * by);
    const float ey = -(uz * bx - ux * bz);
    const float ez = -(ux * by - uy * bx);
    const float pmx = Ux + ex, pmy = Uy + ey, pmz = Uz + ez;
    // 1/gamma- from the hardware reciprocal square root of 1 + |U-|^2
    const float r = __builtin_amdgcn_rsqf(fmaf(pmx, pmx, fmaf(pmy, pmy, fmaf(pmz, pmz, 1.0f))));
    const float tx = bx * r, ty = by * r, tz = bz * r;
    // s = 2t/(1+t^2).  For d = t^2 <= 2^-7 (|t| <= 0.088: more than 35 pushes
    // per local gyration) the series 2(1 - d + d^2 - d^3) is within 2d^4 <
    // 8e-9 of 2/(1+d), below FP32 resolution, and needs no transcendental.  A
    // lane with a larger t takes one Newton step on the hardware reciprocal
    // (~0.5 ulp) instead; the branch is taken only by waves holding such a
    // lane, and each lane's value depends on its own t alone.
    const float d = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
    float f = fmaf(d, fmaf(d, fmaf(d, -2.0f, 2.0f), -2.0f), 2.0f);
    if (__builtin_expect(__ballot(d > 0x1p-7f) != 0, 0)) {
        if (d > 0x1p-7f) {
            const float d1 = d + 1.0f;
            const float r0 = __builtin_amdgcn_rcpf(d1);
            f = 2.0f * fmaf(r0, fmaf(-d1, r0, 1.0f), r0);
        }
    }
    const float sx = tx * f, sy = ty * f, sz = tz * f;
    // p' = p- + p- x t and p+ = p- + p' x s, two fused multiply-adds per component
    const float ppx = fmaf(pmy, tz, fmaf(-pmz, ty, pmx));
    const float ppy = fmaf(pmz, tx, fmaf(-pmx, tz, pmy));
    const float ppz = fmaf(pmx, ty, fmaf(-pmy, tx, pmz));
    const float qx = fmaf(ppy, sz, fmaf(-ppz, sy, pmx));
    const float qy = fmaf(ppz, sx, fmaf(-ppx, sz, pmy));
    const float qz = fmaf(ppx, sy, fmaf(-ppy, sx, pmz));
    const float nx = qx + ex, ny = qy + ey, nz = qz + ez;
    const float dk = (Ux + nx) * ex + (Uy + ny) * ey + (Uz + nz) * ez;
    const double P2n = fma(2.0, static_cast<double>(dk), P2);
    inv_gn = __builtin_amdgcn_rsqf(fmaf(nx, nx, fmaf(ny, ny, fmaf(nz, nz, 1.0f))));
    Ux = nx;
    Uy = ny;
    Uz = nz;
    if (renorm) {
        // rescale to sqrt(P2n): the ratio is 1 + O(1e-6), so sqrt(1+d) = 1 + d/2 - d^2/8
        const float u2 = Ux * Ux + Uy * Uy + Uz * Uz;
        const float dd = static_cast<float>(P2n - static_cast<double>(u2)) * __builtin_amdgcn_rcpf(u2);
        const float sc = fmaf(dd, fmaf(-0.125f, dd, 0.5f), 1.0f);
        Ux *= sc;
        Uy *= sc;
        Uz *= sc;
    }
    P2 = P2n;
}

// Bring an offset that left [0, 1) back into its cell range (periodic wrap
// of the cell index when `periodic`); `moved` is set when the cell changes.
__device__ __forceinline__ void wrap_axis(int32_t &cell, float &off, int32_t nc, bool periodic, bool &moved) {
    if (!(off >= 0.0f && off < 1.0f)) {
        moved = true;
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

// Advance one axis of the split position by u * step cells (one rounding).
__device__ __forceinline__ void move_axis(int32_t &cell, float &off, float u, float step, int32_t nc,
                                          bool periodic, bool &moved) {
    off = fmaf(u, step, off);
    wrap_axis(cell, off, nc, periodic, moved);
}

// One axis of the compensated position after a push that took its
// coordinate x = off + comp out of [0, 1): fold comp into the offset
// (Fast2Sum), move the integer part into the cell (periodic wrap of the cell
// index when `periodic`) and recompute x.  `moved` is set.
__device__ __forceinline__ void cross_axis(int32_t &cell, float &off, float &comp, float &x, int32_t nc,
                                           bool periodic, bool &moved) {
    if (!(x >= 0.0f && x < 1.0f)) {
        moved = true;
        const float t = x;
        comp -= t - off;
        off = t;
        const float fl = floorf(off);
        cell += static_cast<int32_t>(fl);
        off -= fl;
        if (off >= 1.0f) {  // -tiny + 1 rounds to 1: stay put at offset 0
            off -= 1.0f;
            cell += 1;
        }
        if (periodic) {
            // a push moves a particle by less than a box length, so one
            // correction almost always suffices; the division is the fallback
            if (cell < 0) {
                cell += nc;
            } else if (cell >= nc) {
                cell -= nc;
            }
            if (__builtin_expect(cell < 0 || cell >= nc, 0)) {
                cell %= nc;
                if (cell < 0) cell += nc;
            }
        }
        x = off + comp;
    }
}

// x == L on an open box (cell nc, offset 0) as the last cell at offset 1:
// the same point, gathered from the same cell, with a coordinate that keeps
// the particle on the boundary path.
__device__ __forceinline__ void to_last_cell(int32_t &cell, float &off, int32_t nc) {
    if (cell == nc) {
        cell = nc - 1;
        off += 1.0f;
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

// Gather cell of one axis for the polynomial path: the coefficients are
// evaluated at the offset itself, and `bias` (nonzero only at x == L, where
// the last cell is evaluated at t = 1 + off) is folded into them.
__device__ __forceinline__ int32_t gather_cell(int32_t cell, int32_t nc, float &bias) {
    int32_t g = cell;
    bias = 0.0f;
    if (g > nc - 1) {
        bias = static_cast<float>(g - (nc - 1));
        g = nc - 1;
    }
    return g < 0 ? 0 : g;
}

template <class T>
__device__ __forceinline__ void put(void *out, int64_t idx, double v) {
    static_cast<T *>(out)[idx] = static_cast<T>(v);
}

__device__ __forceinline__ float mu_f32(float px, float py, float pz, float bx, float by, float bz,
                                        float qsign) {
    const float p2 = px * px + py * py + pz * pz;
    const float b2 = bx * bx + by * by + bz * bz;
    if (!isfinite(p2) || !isfinite(b2) || p2 == 0.0f || b2 == 0.0f) return NAN;
    // 1/(|p||B|) from the hardware reciprocal square roots (~1 ulp each)
    const float mu = qsign * (px * bx + py * by + pz * bz) * (__builtin_amdgcn_rsqf(p2) * __builtin_amdgcn_rsqf(b2));
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// Save-step output of one particle: state before the push of step si, as
// the reference's *_step copies (:351-356) or mu (pipeline :1349-1374).
__device__ __forceinline__ void write_save(int out_kind, int out_f64, void *out_a, void *out_b,
                                           int64_t slot, int64_t n, int64_t p, bool alive,
                                           const int32_t cell[3], const float off[3], const float comp[3],
                                           const float U[3], const float bt[3],
                                           const SplitGeom &g, const Phys &ph) {
    if (out_kind == PT_OUT_PHASE) {
        for (int c = 0; c < 3; ++c) {
            const double xv = alive ? (static_cast<double>(cell[c]) + static_cast<double>(off[c]) +
                                       static_cast<double>(comp[c])) * g.d[c]
                                    : NAN;
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

// Brick record of node (gi, gj, gk) in 32-bit arithmetic (every record
// index of a brick field fits in 31 bits) and the record steps to its +1 neighbours: 1, 8 and 64
// records inside an 8^3 brick, a jump to the next brick from its last node
// along an axis.
struct CornerRecords {
    uint32_t base, sx, sy, sz;
};
__device__ __forceinline__ CornerRecords corner_records(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk) {
    const uint32_t nb0 = static_cast<uint32_t>(fv.nb[0]), nb1 = static_cast<uint32_t>(fv.nb[1]);
    const uint32_t i = static_cast<uint32_t>(gi), j = static_cast<uint32_t>(gj), k = static_cast<uint32_t>(gk);
    const uint32_t brick = ((k >> 3) * nb1 + (j >> 3)) * nb0 + (i >> 3);
    CornerRecords r;
    r.base = (brick << 9) | ((k & 7) << 6) | ((j & 7) << 3) | (i & 7);
    r.sx = (i & 7) == 7 ? 512u - 7u : 1u;
    r.sy = (j & 7) == 7 ? nb0 * 512u - 56u : 8u;
    r.sz = (k & 7) == 7 ? nb0 * nb1 * 512u - 448u : 64u;
    return r;
}

// Corner values of the six components at the 8 nodes of cell (gi,gj,gk),
// corner index ci = dx + 2 dy + 4 dz, in storage units.
template <int STORE>
__device__ __forceinline__ void load_corners(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk,
                                             float s[6][8]) {
    if constexpr (STORE == PT_STORE_F32_PLANAR) {
#pragma unroll
        for (int ci = 0; ci < 8; ++ci) {
            const int64_t i = gi + (ci & 1), j = gj + ((ci >> 1) & 1), k = gk + (ci >> 2);
            const int64_t nn = fv.n[0] * fv.n[1] * fv.n[2];
            const int64_t idx = i + fv.n[0] * (j + fv.n[1] * k);
            const float *F = static_cast<const float *>(fv.data);
#pragma unroll
            for (int c = 0; c < 6; ++c) s[c][ci] = F[c * nn + idx];
        }
    } else {
        // Record of corner (0,0,0), then per-axis steps to the +1 neighbour:
        // inside an 8^3 brick the steps are 1, 8 and 64 records; from the last
        // node of a brick they jump to the next brick along that axis.  Every
        // record index fits in 32 bits (< 2^31 / 16 bytes for any field the
        // library accepts).
        const CornerRecords cr = corner_records(fv, gi, gj, gk);
        const uint32_t rec[8] = {cr.base, cr.base + cr.sx, cr.base + cr.sy, cr.base + cr.sy + cr.sx,
                                 cr.base + cr.sz, cr.base + cr.sz + cr.sx, cr.base + cr.sz + cr.sy,
                                 cr.base + cr.sz + cr.sy + cr.sx};
#pragma unroll
        for (int ci = 0; ci < 8; ++ci) {
            if constexpr (STORE == PT_STORE_F32_BRICK) {
                const float4 *R = static_cast<const float4 *>(fv.data) + static_cast<uint64_t>(rec[ci]) * 2u;
                const float4 a = R[0];
                const float4 b = R[1];
                s[0][ci] = a.x;
                s[1][ci] = a.y;
                s[2][ci] = a.z;
                s[3][ci] = a.w;
                s[4][ci] = b.x;
                s[5][ci] = b.y;
            } else {
                const uint4 w = static_cast<const uint4 *>(fv.data)[rec[ci]];
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
// decoded (stored * scale) and multiplied by the physical factor.
__device__ __forceinline__ void coeffs(const float s[8], float scale, float a[8]) {
    const float dx00 = s[1] - s[0], dx10 = s[3] - s[2], dx01 = s[5] - s[4], dx11 = s[7] - s[6];
    const float dy0 = s[2] - s[0], dy1 = s[6] - s[4];
    a[0] = s[0] * scale;
    a[1] = dx00 * scale;
    a[2] = dy0 * scale;
    a[3] = (s[4] - s[0]) * scale;
    a[4] = (dx10 - dx00) * scale;
    a[5] = (dx01 - dx00) * scale;
    a[6] = (dy1 - dy0) * scale;
    a[7] = ((dx11 - dx01) - (dx10 - dx00)) * scale;
}

// Coefficients of f(x + bx, y + by, z + bz), one axis at a time.
__device__ __forceinline__ void shift_coeffs(float a[8], float bx, float by, float bz) {
    a[0] = fmaf(a[1], bx, a[0]);
    a[2] = fmaf(a[4], bx, a[2]);
    a[3] = fmaf(a[5], bx, a[3]);
    a[6] = fmaf(a[7], bx, a[6]);
    a[0] = fmaf(a[2], by, a[0]);
    a[1] = fmaf(a[4], by, a[1]);
    a[3] = fmaf(a[6], by, a[3]);
    a[5] = fmaf(a[7], by, a[5]);
    a[0] = fmaf(a[3], bz, a[0]);
    a[1] = fmaf(a[5], bz, a[1]);
    a[2] = fmaf(a[6], bz, a[2]);
    a[4] = fmaf(a[7], bz, a[4]);
}

// Add the coarse part of one component over fine cell (gi, gj, gk) of a
// 16-bit brick field to its residual polynomial a: the coarse cell's
// polynomial (fine-node units from its first node, in units of the residual
// scale) moved to the cell's origin, times `scale` (the residual scale times
// the physical factor).  The cell lies inside coarse cell (gi >> 2, gj >> 2,
// gk >> 2).  The table index counts 16-byte units in 32 bits: a brick field
// has fewer than 2^31 nodes, so fewer than 2^26 coarse cells.
__device__ __forceinline__ void add_coarse(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk, int comp,
                                           float scale, float a[8]) {
    const uint32_t cell = static_cast<uint32_t>(
        (gi >> kCoarseShift) + fv.cn[0] * ((gj >> kCoarseShift) + fv.cn[1] * (gk >> kCoarseShift)));
    const float4 *src = reinterpret_cast<const float4 *>(fv.coarse) + (cell * 6u + static_cast<uint32_t>(comp)) * 2u;
    const float4 lo = src[0], hi = src[1];
    float A[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    shift_coeffs(A, static_cast<float>(gi & (kCoarseWidth - 1)), static_cast<float>(gj & (kCoarseWidth - 1)),
                 static_cast<float>(gk & (kCoarseWidth - 1)));
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = fmaf(A[q], scale, a[q]);
}

// Cell polynomial of one component from its 8 corner values in storage
// units times `scale`.  16-bit bricks then add their coarse part.  kSerial:
// its loads are issued only after the residual coefficients are formed (the
// barrier keeps the scheduler from holding both sets of loads in flight);
// the per-lane reload of all six components needs that to stay within the
// register budget, while a helper lane of the cooperative pass (one
// component) keeps both in flight, one memory round trip instead of two.
// Scheduling only: the arithmetic is the same either way.
template <int STORE, bool kSerial = true>
__device__ __forceinline__ void cell_coeffs(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk, int comp,
                                            const float s[8], float scale, float a[8]) {
    coeffs(s, scale, a);
    if constexpr (kHierarchical<STORE>) {
        if constexpr (kSerial) __builtin_amdgcn_sched_barrier(0);
        add_coarse(fv, gi, gj, gk, comp, scale, a);
    }
}

__device__ __forceinline__ float eval_poly(const float a[8], float x, float y, float z) {
    const float z0 = fmaf(a[3], z, a[0]);
    const float z1 = fmaf(a[5], z, a[1]);
    const float z2 = fmaf(a[6], z, a[2]);
    const float z3 = fmaf(a[7], z, a[4]);
    return fmaf(fmaf(z3, x, z2), y, fmaf(z1, x, z0));
}


// The 8 corner values of one component of cell (gi,gj,gk), in storage
// units (corner ci = dx + 2 dy + 4 dz).
template <int STORE>
__device__ __forceinline__ void load_component_corners(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk,
                                                       int comp, float s[8]) {
    if constexpr (STORE == PT_STORE_F32_PLANAR) {
        const int64_t nn = fv.n[0] * fv.n[1] * fv.n[2];
        const float *F = static_cast<const float *>(fv.data) + comp * nn;
#pragma unroll
        for (int ci = 0; ci < 8; ++ci) {
            const int64_t i = gi + (ci & 1), j = gj + ((ci >> 1) & 1), k = gk + (ci >> 2);
            s[ci] = F[i + fv.n[0] * (j + fv.n[1] * k)];
        }
    } else {
        const CornerRecords cr = corner_records(fv, gi, gj, gk);
        // This component's half of a 32-bit word, taken per lane without a
        // select between the two halves (with a select the compiler waited
        // on each of the eight loads in turn): FP16 by a shift, BF16 by one
        // byte permute that moves the half to the top and zeroes the rest.
        const uint32_t sh = (comp & 1) ? 16u : 0u;
        const uint32_t sel = (comp & 1) ? 0x03020c0cu : 0x01000c0cu;
        if constexpr (STORE == PT_STORE_BF16_BRICK) {
            // the eight loads first, then the decodes: the scheduler keeps
            // them apart, so all eight are in flight together
            uint32_t w[8];
            const uint32_t *src = static_cast<const uint32_t *>(fv.data) + (comp >> 1);
#pragma unroll
            for (int ci = 0; ci < 8; ++ci) {
                const uint32_t rec = cr.base + ((ci & 1) ? cr.sx : 0u) + ((ci & 2) ? cr.sy : 0u) + ((ci & 4) ? cr.sz : 0u);
                w[ci] = src[static_cast<uint64_t>(rec) * 4u];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int ci = 0; ci < 8; ++ci) s[ci] = __uint_as_float(__builtin_amdgcn_perm(w[ci], w[ci], sel));
        } else {
#pragma unroll
            for (int ci = 0; ci < 8; ++ci) {
                const uint32_t rec = cr.base + ((ci & 1) ? cr.sx : 0u) + ((ci & 2) ? cr.sy : 0u) + ((ci & 4) ? cr.sz : 0u);
                if constexpr (STORE == PT_STORE_F32_BRICK) {
                    s[ci] = (static_cast<const float *>(fv.data) + comp)[static_cast<uint64_t>(rec) * 8u];
                } else {
                    const uint32_t w = (static_cast<const uint32_t *>(fv.data) + (comp >> 1))[static_cast<uint64_t>(rec) * 4u];
                    s[ci] = f16_bits_to_float(static_cast<uint16_t>((w >> sh) & 0xffffu));
                }
            }
        }
    }
}

// Wave-cooperative coefficient reload.  A lane that changed cell posts its
// gather cell in the wave's LDS area; six helper lanes per posted cell (one
// per field component) load that component's 8 corners and compute its 8
// polynomial coefficients; the lane then reads its 48 coefficients back.
// Up to kReloadBatch cells are served per pass, so a wave pays one pass when
// a few of its lanes change cell, instead of every lane's share of a full
// per-lane reload.  Must be called with all 64 lanes of the wave active.
constexpr int kReloadBatch = 10;
struct ReloadLds {
    int4 slot[kReloadBatch];
    float coef[kReloadBatch][6][8];
    float scale[6];  // per component: decode scale (times the physical factor unless 16-bit)
    float fac[6];    // per component: physical factor (16-bit bricks)
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane index within the wave, and the number of lanes below it whose bit
// is set in `mask` (hardware bit counts, no per-lane constants kept).
__device__ __forceinline__ int lane_id() {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
}
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
}

// The per-component scales and factors the helper lanes apply, written once
// into the wave's LDS area (they are read only on a reload pass, so no lane
// keeps them in a register across the step loop).
template <int STORE>
__device__ __forceinline__ void reload_init(const FieldView &fv, const Phys &ph, ReloadLds *lds) {
    const int lane = lane_id();
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        if (lane == c) {
            lds->scale[c] = kHierarchical<STORE> ? fv.scale[c] : fv.scale[c] * ph.fac[c];
            lds->fac[c] = ph.fac[c];
        }
    }
    wave_lds_sync();
}

// A particle's coefficients are the cell's field-unit coefficients (what a
// cell table holds) times the component's physical factor.  FP32 storages
// have scale 1, so folding the factor into the decode scale gives the same
// product; 16-bit bricks form the field-unit values (residual plus coarse
// part) first and multiply after.  Every reload path, with or without a cell
// table, therefore gives the same values bit for bit.
__device__ __forceinline__ void apply_factor(float a[8], float fac) {
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] *= fac;
}

// The 48 coefficients of gather cell (gi, gj, gk) from the field's cell table.
__device__ __forceinline__ void table_coeffs(const FieldView &fv, int32_t gi, int32_t gj, int32_t gk,
                                             const Phys &ph, float a[6][8]) {
    const float4 *src = reinterpret_cast<const float4 *>(fv.cells) + cell_record_index(fv.cb, gi, gj, gk) * 12;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float4 v = src[q];
        float *d = &a[q >> 1][(q & 1) * 4];
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) apply_factor(a[c], ph.fac[c]);
}

// Per-lane reload of the coefficients of the particle's own cell, all loads
// in flight at once (used before the step loop, where every lane reloads).
template <int STORE>
__device__ __forceinline__ void lane_reload(const int32_t cell[3], const FieldView &fv, const SplitGeom &g,
                                            const Phys &ph, float a[6][8]) {
    float b[3];
    const int32_t gi = gather_cell(cell[0], g.nc[0], b[0]);
    const int32_t gj = gather_cell(cell[1], g.nc[1], b[1]);
    const int32_t gk = gather_cell(cell[2], g.nc[2], b[2]);
    if constexpr (STORE == kStoreCells) {
        table_coeffs(fv, gi, gj, gk, ph, a);
    } else {
        float sv[6][8];
        load_corners<STORE>(fv, gi, gj, gk, sv);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if constexpr (kHierarchical<STORE>) {
                cell_coeffs<STORE>(fv, gi, gj, gk, c, sv[c], fv.scale[c], a[c]);
                apply_factor(a[c], ph.fac[c]);
            } else {
                cell_coeffs<STORE>(fv, gi, gj, gk, c, sv[c], fv.scale[c] * ph.fac[c], a[c]);
            }
        }
    }
    if (b[0] != 0.0f || b[1] != 0.0f || b[2] != 0.0f) {
#pragma unroll
        for (int c = 0; c < 6; ++c) shift_coeffs(a[c], b[0], b[1], b[2]);
    }
}

template <int STORE>
__device__ __forceinline__ void coop_reload(int &need, const int32_t cell[3], const FieldView &fv,
                                            const SplitGeom &g, const Phys &ph, ReloadLds *lds, float a[6][8]) {
    if constexpr (STORE == kStoreCells) {
        // a cell table needs no helpers: each lane that changed cell loads
        // its own 48 coefficients
        if (__ballot(need != 0)) {
            if (need) {
                lane_reload<STORE>(cell, fv, g, ph, a);
                need = 0;
            }
        }
        return;
    }
    uint64_t pending = __ballot(need != 0);
    while (pending) {
        const int rank = lanes_below(pending);
        const bool mine = need != 0 && rank < kReloadBatch;
        const int m = __popcll(pending) < kReloadBatch ? __popcll(pending) : kReloadBatch;
        if (mine) {
            float b[3];
            const int32_t gi = gather_cell(cell[0], g.nc[0], b[0]);
            const int32_t gj = gather_cell(cell[1], g.nc[1], b[1]);
            const int32_t gk = gather_cell(cell[2], g.nc[2], b[2]);
            // a live particle is at most one cell past the last gather cell (x == L)
            const int bias = (b[0] != 0.0f ? 1 : 0) | (b[1] != 0.0f ? 2 : 0) | (b[2] != 0.0f ? 4 : 0);
            lds->slot[rank] = make_int4(gi, gj, gk, bias);
        }
        wave_lds_sync();
        // helper role: lane = 6 entry + comp serves component comp of
        // posted cell `entry` ((lane * 43) >> 8 == lane / 6 for lane < 64)
        const int lane = lane_id();
        const int entry = (lane * 43) >> 8;
        const int comp = lane - 6 * entry;
        if (entry < m) {
            const int4 t = lds->slot[entry];
            float sv[8], c8[8];
            load_component_corners<STORE>(fv, t.x, t.y, t.z, comp, sv);
            cell_coeffs<STORE, false>(fv, t.x, t.y, t.z, comp, sv, lds->scale[comp], c8);
            if constexpr (kHierarchical<STORE>) apply_factor(c8, lds->fac[comp]);
            if (t.w) shift_coeffs(c8, (t.w & 1) ? 1.0f : 0.0f, (t.w & 2) ? 1.0f : 0.0f, (t.w & 4) ? 1.0f : 0.0f);
            float4 *dst = reinterpret_cast<float4 *>(lds->coef[entry][comp]);
            dst[0] = make_float4(c8[0], c8[1], c8[2], c8[3]);
            dst[1] = make_float4(c8[4], c8[5], c8[6], c8[7]);
        }
        wave_lds_sync();
        if (mine) {
            const float4 *src = reinterpret_cast<const float4 *>(lds->coef[rank]);
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const float4 v = src[q];
                float *d = &a[q >> 1][(q & 1) * 4];
                d[0] = v.x;
                d[1] = v.y;
                d[2] = v.z;
                d[3] = v.w;
            }
            need = 0;
        }
        pending &= ~__ballot(mine);
        wave_lds_sync();  // the next pass overwrites the slots and coefficients
    }
}

// Occupancy of the kernels that run the step loop (the single-campaign
// pusher and the serving tick): four waves per SIMD (at most 128 VGPRs) on
// stored nodes, the budget at which no storage spills.  With five (at most
// 102 VGPRs) the compiler kept launch constants of the FP32 storages in
// scratch (a 64-byte spill per lane written at every launch) and, for 16-bit
// bricks, values the push itself reads (a 16-byte reload every push);
// measured per 200-push launch, five waves against four: 4,194,304 protons
// at 100 GeV, FP32 bricks 2.10 against 2.10 ms, BF16 2.6 against 2.2 ms,
// FP16 3.6 against 2.2 ms; 1,048,576 at 100 TeV, FP32 bricks 1.74 against
// 1.59 ms.  A cell table keeps five waves: its per-lane record loads gain
// more from the fifth wave than its spill of launch constants costs (10 TeV:
// 0.80 against 0.84 ms).
template <int STORE>
constexpr int kPushWaves = STORE == kStoreCells ? 5 : 4;

// Cadence of one launch: steps si0 .. si0+K-1; a save happens where
// (si-1) % stride == 0.  phase0 = (si0-1) % stride and slot0 = the index of
// the first save at or after si0 are computed once on the host, so the loop
// keeps 32-bit counters instead of dividing 64-bit step indices.
struct Cadence {
    int64_t si0, total_steps;
    int32_t K, stride, phase0, want_save;
    int64_t slot0;
};

__host__ inline Cadence make_cadence(int64_t si0, int K, int64_t total_steps, int64_t stride, bool want_save,
                                     int64_t save_base) {
    Cadence c;
    c.si0 = si0;
    c.total_steps = total_steps;
    c.K = K;
    c.stride = static_cast<int32_t>(stride);
    c.phase0 = static_cast<int32_t>((si0 - 1) % stride);
    c.want_save = want_save ? 1 : 0;
    c.slot0 = (si0 - 1 + stride - 1) / stride - save_base;
    return c;
}

// The production step loop of one lane: K steps with state and the current
// cell's coefficients in registers.  The coefficients are reloaded only after
// the particle changes cell, so a push costs the six polynomial evaluations,
// the Boris step and the position update.  save(slot, alive, cell, off,
// comp, U, P2, v) is called at every cadence step (before the push, as the
// reference's *_step copies); die(si) when an open boundary removes the
// particle.
//
// Position: cell + off + comp.  Each push adds its step u * k / gamma to the
// small term comp (one rounding at the scale of comp, not of the offset), and
// the polynomial is evaluated at x = off + comp.  Every fourth global step
// (and whenever x leaves [0, 1)) comp is folded into the offset with
// compensated summation (Fast2Sum: off = x, comp keeps what that addition
// dropped), so a step far below the offset's ulp (6e-8 of a cell near 1)
// still accumulates at its full value instead of rounding the same way push
// after push.  The fold cadence follows the global step index, so the result
// does not depend on how the steps are split into launches.
//
// All 64 lanes of the wave must run the loop (lanes without a particle with
// alive = false): the coefficient reload is wave-cooperative (coop_reload)
// and uses the wave's `lds` area.
template <int STORE, bool CUBIC, class SaveFn, class DieFn>
__device__ __forceinline__ void run_lane(int32_t cell[3], float off[3], float comp[3], float U[3], double &P2,
                                         bool &alive,
                                         const FieldView &fv, const SplitGeom &g, const Phys &ph,
                                         const Cadence &cad, int periodic, ReloadLds *lds, SaveFn save, DieFn die) {
    // A particle on an upper face of an open box (x == L: cell nc, offset
    // 0) is held as the last cell at offset 1, so its coordinate is not
    // below 1 and every push takes the boundary path until it moves inwards
    // (or leaves).
    if (!periodic) {
#pragma unroll
        for (int c = 0; c < 3; ++c) to_last_cell(cell[c], off[c], g.nc[c]);
    }
    float a[6][8];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int q = 0; q < 8; ++q) a[c][q] = 0.0f;
    if (alive) lane_reload<STORE>(cell, fv, g, ph, a);
    // Lane flags as integers, compared against zero where they are used:
    // live = the particle is alive; reload = it changed cell in the previous
    // push (never set on a dead lane).
    int live = alive ? 1 : 0;
    int reload = 0;
    float x[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = off[c] + comp[c];
    const int32_t si_mod4 = static_cast<int32_t>(cad.si0 & 3);
    reload_init<STORE>(fv, ph, lds);
    int32_t phase = cad.phase0;
    int64_t slot = cad.slot0;
    // pushes happen while si < total_steps: at most n_push of the K steps
    const int64_t room = cad.total_steps - cad.si0;
    const int32_t n_push = room <= 0 ? 0 : (room < cad.K ? static_cast<int32_t>(room) : cad.K);
    auto fold = [&](int c) {
        const float t = x[c];
        comp[c] -= t - off[c];
        off[c] = t;
    };
    for (int32_t k = 0; k < cad.K; ++k) {
        const bool do_save = cad.want_save && phase == 0;
        phase = phase + 1 == cad.stride ? 0 : phase + 1;
        const bool push = live != 0 && k < n_push;
        coop_reload<STORE>(reload, cell, fv, g, ph, lds, a);  // every lane of the wave, converged here
        if (!do_save && !push) continue;
        // dead lanes evaluate stale coefficients only to keep the loop
        // uniform; save() writes NaN for them
        float v[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) v[c] = eval_poly(a[c], x[0], x[1], x[2]);
        if (do_save) {
            save(slot, live != 0, cell, off, comp, U, P2, v);
            ++slot;
        }
        if (!push) continue;
        // |U| is rescaled after the pushes of global steps si = 4, 8, ... and
        // before each save (phase == 0: the next step saves), so every output
        // carries the rescaled magnitude and the result does not depend on
        // how the steps are split into launches
        const bool fourth = ((si_mod4 + k) & 3) == 0;
        float inv_gn;
        boris_split(U[0], U[1], U[2], P2, v[0], v[1], v[2], v[3], v[4], v[5], inv_gn, fourth | (phase == 0));
        // With cubic cells (CUBIC: g.k equal on every axis) the step factor
        // is one product, the same value each axis would compute.
        if constexpr (CUBIC) {
            const float w = g.k[0] * inv_gn;
#pragma unroll
            for (int c = 0; c < 3; ++c) comp[c] = fmaf(U[c], w, comp[c]);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) comp[c] = fmaf(U[c], g.k[c] * inv_gn, comp[c]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = off[c] + comp[c];
        if (fourth) {
            // a branch, not selects: the empty asm keeps the compiler from
            // flattening the fold into every push
            asm volatile("");
#pragma unroll
            for (int c = 0; c < 3; ++c) fold(c);
        }
        // every coordinate in [0, 1) <=> every bit pattern below that of
        // 1.0f (negative values, -0 and NaN compare above it)
        const uint32_t top = max(max(__float_as_uint(x[0]), __float_as_uint(x[1])), __float_as_uint(x[2]));
        if (__builtin_expect(!(top < 0x3f800000u), 0)) {
            bool moved = false;
#pragma unroll
            for (int c = 0; c < 3; ++c) cross_axis(cell[c], off[c], comp[c], x[c], g.nc[c], periodic, moved);
            reload |= moved ? 1 : 0;
            if (!periodic) {
                if (outside(cell[0], off[0], g.nc[0]) | outside(cell[1], off[1], g.nc[1]) |
                    outside(cell[2], off[2], g.nc[2])) {
                    die(cad.si0 + k);
                    live = 0;
                    reload = 0;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    to_last_cell(cell[c], off[c], g.nc[c]);
                    x[c] = off[c] + comp[c];
                }
            }
        }
    }
    alive = live != 0;
}

}  // namespace core
}  // namespace pt
