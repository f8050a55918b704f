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

#include "pt_push_core.hpp"

namespace pt {
namespace {

using namespace core;

// ----------------------------------------------------------- plain ---------

__global__ void __launch_bounds__(256)
k_push_plain(SplitState s, uint8_t *__restrict__ alive_arr, const float *__restrict__ F,
             int64_t nx, int64_t ny, int64_t nz, SplitGeom g, Phys ph, int64_t n, int64_t si,
             int64_t total_steps, int64_t stride, int64_t save_base, int out_kind, int out_f64,
             void *out_a, void *out_b, int periodic, int64_t first_step,
             unsigned long long *deaths, const int32_t *__restrict__ perm) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t col = perm ? perm[p] : p;
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
    double P2 = s.P2[p] * static_cast<double>(ph.mr) * static_cast<double>(ph.mr);
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
        // the plain path adds each step to the FP32 offset directly, as the
        // reference adds it to the position; comp stays the residual of the
        // initial position
        float comp[3];
        for (int c = 0; c < 3; ++c) comp[c] = s.comp[c][p];
        write_save(out_kind, out_f64, out_a, out_b, slot, n, col, alive, cell, off, comp, U, v, g, ph);
    }
    if (!push) return;
    float inv_gn;
    bool moved = false;
    boris_split(U[0], U[1], U[2], P2, v[0], v[1], v[2], v[3], v[4], v[5], inv_gn);
    for (int c = 0; c < 3; ++c) move_axis(cell[c], off[c], U[c], g.k[c] * inv_gn, g.nc[c], periodic, moved);
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
    s.P2[p] = P2 / (static_cast<double>(ph.mr) * static_cast<double>(ph.mr));
}

// ------------------------------------------------------------ fast ---------

// The production kernel's body.
template <int STORE, bool CUBIC>
__device__ __forceinline__ void push_fast(SplitState s, uint8_t *__restrict__ alive_arr, const FieldView &fv,
                                          const SplitGeom &g, const Phys &ph, int64_t n, const Cadence &cad,
                                          int out_kind, int out_f64, void *out_a, void *out_b, int periodic,
                                          int64_t first_step, unsigned long long *deaths,
                                          const int32_t *__restrict__ perm) {
    __shared__ ReloadLds lds[4];  // one per wave of the 256-lane block
    const int64_t nblk = gridDim.x;
    const int64_t p = xcd_block(blockIdx.x, nblk) * blockDim.x + threadIdx.x;
    // Lanes past the last particle stay in the loop as dead helpers of the
    // wave-cooperative reload; a wave without any particle leaves at once.
    const bool valid = p < n;
    if (__ballot(valid) == 0) return;
    const int64_t q = valid ? p : 0;  // in-range index for their loads
    const int64_t col = perm ? perm[q] : q;
    bool alive = valid && alive_arr[q] != 0;
    int32_t cell[3];
    float off[3], comp[3], U[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cell[c] = s.cell[c][q];
        off[c] = s.off[c][q];
        comp[c] = s.comp[c][q];
        U[c] = s.P[c][q] * ph.mr;
    }
    const double mr2 = static_cast<double>(ph.mr) * static_cast<double>(ph.mr);
    double P2 = s.P2[q] * mr2;
    run_lane<STORE, CUBIC>(
        cell, off, comp, U, P2, alive, fv, g, ph, cad, periodic, &lds[threadIdx.x >> 6],
        [&](int64_t slot, bool alv, const int32_t *ce, const float *of, const float *cp, const float *u, double,
            const float *v) {
            if (valid) write_save(out_kind, out_f64, out_a, out_b, slot, n, col, alv, ce, of, cp, u, v, g, ph);
        },
        [&](int64_t si) { atomicAdd(&deaths[si - first_step], 1ull); });
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.cell[c][p] = cell[c];
        s.off[c][p] = off[c];
        s.comp[c][p] = comp[c];
        s.P[c][p] = U[c] / ph.mr;
    }
    s.P2[p] = P2 / mr2;
    alive_arr[p] = alive ? 1 : 0;
}

#define PT_FAST_PARAMS                                                                                     \
    SplitState s, uint8_t *__restrict__ alive_arr, FieldView fv, SplitGeom g, Phys ph, int64_t n, Cadence cad, \
        int out_kind, int out_f64, void *out_a, void *out_b, int periodic, int64_t first_step,             \
        unsigned long long *deaths, const int32_t *__restrict__ perm
#define PT_FAST_ARGS s, alive_arr, fv, g, ph, n, cad, out_kind, out_f64, out_a, out_b, periodic, first_step, deaths, perm

// Waves per SIMD: kPushWaves (pt_push_core.hpp).
template <int STORE, bool CUBIC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPushWaves<STORE>, 8)))
k_push_fast(PT_FAST_PARAMS) {
    push_fast<STORE, CUBIC>(PT_FAST_ARGS);
}
#undef PT_FAST_PARAMS
#undef PT_FAST_ARGS

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

// One k_push_fast launch; cubic cells (the same c dt / dx on every axis)
// take the variant that forms the position step factor once per push.
template <int STORE, class... Args>
void launch_fast(const SplitGeom &g, unsigned blocks, hipStream_t stream, Args... args) {
    if (g.k[0] == g.k[1] && g.k[0] == g.k[2])
        k_push_fast<STORE, true><<<blocks, 256, 0, stream>>>(args...);
    else
        k_push_fast<STORE, false><<<blocks, 256, 0, stream>>>(args...);
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
            a.out_a, a.out_b, e->periodic, a.death_base, reinterpret_cast<unsigned long long *>(e->deaths), e->perm);
        PT_HIP_CHECK(hipGetLastError());
        ++launches;
    }
    return launches;
}

int64_t launch_push_fast(pt_ens *e, const pt_field *f, const PushArgs &a) {
    SplitGeom g;
    Phys ph;
    make_geom_phys(f, a, g, ph);
    const FieldView fv = make_field_view(f);
    const int K = a.steps_per_launch > 0 ? a.steps_per_launch : 128;
    const unsigned blocks = static_cast<unsigned>((e->n + 255) / 256);
    int64_t launches = 0;
    for (int64_t done = 0; done < a.n_steps; done += K) {
        const int k = static_cast<int>(std::min<int64_t>(K, a.n_steps - done));
        auto *dl = reinterpret_cast<unsigned long long *>(e->deaths);
        const int64_t si0 = a.first_step + done;
        const Cadence cad = make_cadence(si0, k, a.total_steps, a.stride, a.out_kind != PT_OUT_NONE, a.save_base);
        if (f->cells) {
            launch_fast<kStoreCells>(g, blocks, e->ctx->stream, split_state(e), e->alive, fv, g, ph, e->n, cad,
                                     a.out_kind, a.out_f64, a.out_a, a.out_b, e->periodic, a.death_base, dl, e->perm);
            PT_HIP_CHECK(hipGetLastError());
            ++launches;
            continue;
        }
        switch (f->storage) {
#define PT_LAUNCH_FAST(ST)                                                                                  \
    case ST:                                                                                                \
        launch_fast<ST>(g, blocks, e->ctx->stream, split_state(e), e->alive, fv, g, ph, e->n, cad, a.out_kind, \
                        a.out_f64, a.out_a, a.out_b, e->periodic, a.death_base, dl, e->perm);               \
        break;
            PT_LAUNCH_FAST(PT_STORE_F32_PLANAR)
            PT_LAUNCH_FAST(PT_STORE_F32_BRICK)
            PT_LAUNCH_FAST(PT_STORE_BF16_BRICK)
            PT_LAUNCH_FAST(PT_STORE_F16_BRICK)
#undef PT_LAUNCH_FAST
            default:
                throw Error{PT_ERR_INVALID, "PT_PUSH_FAST: unsupported field storage"};
        }
        PT_HIP_CHECK(hipGetLastError());
        ++launches;
    }
    return launches;
}

}  // namespace pt
