// FP64 reference-structured pusher (PT_PUSH_REF_F64): the parity path.
//
// Follows advance_particles_kernel! (phase_space_gpu_runner.jl:330-391) with
// trilinear_sample (:247-276), cross3 (:278-280), boris_push (:282-323) and
// wrap_periodic_coordinate (:325-328) in the reference's operation order.
// This file is compiled with -ffp-contract=off so no multiply-add is fused,
// as the reference's FP64 arithmetic is written.  One thread per particle;
// the particle's state stays in registers across the K steps of a launch and
// the save-step outputs (phase space, or mu as reconstruct_mu_step_kernel!,
// multimode_multienergy_cache_pipeline.jl:1349-1374) are written at cadence.
#include <algorithm>

#include "pt_internal.hpp"

namespace pt {
namespace {

struct GeomD {
    double lo[3], hi[3], d[3];
    int64_t n[3];
};

__device__ __forceinline__ int64_t clamp_floor(double f, int64_t lo, int64_t hi) {
    double fl = floor(f);
    if (fl < static_cast<double>(lo)) return lo;
    if (fl > static_cast<double>(hi)) return hi;
    return static_cast<int64_t>(fl);
}

// trilinear_sample, phase_space_gpu_runner.jl:247-276 (1-based indices kept).
__device__ double trilinear_ref(const double *__restrict__ arr, double x, double y, double z,
                                const GeomD &g) {
    const double fx = (x - g.lo[0]) / g.d[0] + 1.0;
    const double fy = (y - g.lo[1]) / g.d[1] + 1.0;
    const double fz = (z - g.lo[2]) / g.d[2] + 1.0;
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t i0 = clamp_floor(fx, 1, nx - 1);
    const int64_t j0 = clamp_floor(fy, 1, ny - 1);
    const int64_t k0 = clamp_floor(fz, 1, nz - 1);
    const double tx = fx - static_cast<double>(i0);
    const double ty = fy - static_cast<double>(j0);
    const double tz = fz - static_cast<double>(k0);
    const int64_t base = (i0 - 1) + nx * ((j0 - 1) + ny * (k0 - 1));
    const int64_t sy = nx, sz = nx * ny;
    const double c000 = arr[base];
    const double c100 = arr[base + 1];
    const double c010 = arr[base + sy];
    const double c110 = arr[base + 1 + sy];
    const double c001 = arr[base + sz];
    const double c101 = arr[base + 1 + sz];
    const double c011 = arr[base + sy + sz];
    const double c111 = arr[base + 1 + sy + sz];
    const double c00 = c000 * (1 - tx) + c100 * tx;
    const double c10 = c010 * (1 - tx) + c110 * tx;
    const double c01 = c001 * (1 - tx) + c101 * tx;
    const double c11 = c011 * (1 - tx) + c111 * tx;
    const double c0 = c00 * (1 - ty) + c10 * ty;
    const double c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

__device__ __forceinline__ double wrap_ref(double x, double xmin, double xmax) {
    const double width = xmax - xmin;
    return x - floor((x - xmin) / width) * width;
}

// pitch_mu_from_pB, compute_delta_mu2_curve.jl:101-110.
__device__ __forceinline__ double mu_ref(double px, double py, double pz, double bx, double by,
                                         double bz) {
    const double pn = sqrt(px * px + py * py + pz * pz);
    const double bn = sqrt(bx * bx + by * by + bz * bz);
    if (!isfinite(pn) || !isfinite(bn) || pn == 0.0 || bn == 0.0) return NAN;
    double mu = (px * bx + py * by + pz * bz) / (pn * bn);
    return fmin(fmax(mu, -1.0), 1.0);
}

template <class T>
__device__ __forceinline__ void put(void *out, int64_t idx, double v) {
    static_cast<T *>(out)[idx] = static_cast<T>(v);
}

__global__ void __launch_bounds__(256)
k_push_ref_f64(RefState s, uint8_t *__restrict__ alive_arr, const double *__restrict__ F, GeomD g,
               int64_t n, int64_t si0, int K, int64_t total_steps, int64_t stride, int64_t save_base,
               int out_kind, int out_f64, void *out_a, void *out_b, double dt, double qm, double mp,
               int periodic, int64_t first_step, unsigned long long *deaths) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t nn = g.n[0] * g.n[1] * g.n[2];
    const double *Bx = F, *By = F + nn, *Bz = F + 2 * nn;
    const double *Vx = F + 3 * nn, *Vy = F + 4 * nn, *Vz = F + 5 * nn;
    bool alive = alive_arr[p] != 0;
    double x = s.x[0][p], y = s.x[1][p], z = s.x[2][p];
    double px = s.p[0][p], py = s.p[1][p], pz = s.p[2][p];
    const double cl = kC;
    for (int k = 0; k < K; ++k) {
        const int64_t si = si0 + k;
        if (out_kind != PT_OUT_NONE && ((si - 1) % stride) == 0) {
            const int64_t slot = (si - 1) / stride - save_base;
            if (out_kind == PT_OUT_PHASE) {
                const double vals[6] = {alive ? x : NAN,  alive ? y : NAN,  alive ? z : NAN,
                                        alive ? px : NAN, alive ? py : NAN, alive ? pz : NAN};
                for (int c = 0; c < 3; ++c) {
                    const int64_t idx = (slot * 3 + c) * n + p;
                    if (out_f64) {
                        put<double>(out_a, idx, vals[c]);
                        put<double>(out_b, idx, vals[3 + c]);
                    } else {
                        put<float>(out_a, idx, vals[c]);
                        put<float>(out_b, idx, vals[3 + c]);
                    }
                }
            } else {
                double mu = NAN;
                if (alive && g.lo[0] <= x && x <= g.hi[0] && g.lo[1] <= y && y <= g.hi[1] &&
                    g.lo[2] <= z && z <= g.hi[2]) {
                    const double bx = trilinear_ref(Bx, x, y, z, g);
                    const double by = trilinear_ref(By, x, y, z, g);
                    const double bz = trilinear_ref(Bz, x, y, z, g);
                    mu = mu_ref(px, py, pz, bx, by, bz);
                }
                if (out_f64)
                    put<double>(out_a, slot * n + p, mu);
                else
                    put<float>(out_a, slot * n + p, mu);
            }
        }
        if (!alive || si >= total_steps) continue;
        // advance_particles_kernel! push branch, phase_space_gpu_runner.jl:358-388
        const double bx = trilinear_ref(Bx, x, y, z, g);
        const double by = trilinear_ref(By, x, y, z, g);
        const double bz = trilinear_ref(Bz, x, y, z, g);
        const double vfx = trilinear_ref(Vx, x, y, z, g);
        const double vfy = trilinear_ref(Vy, x, y, z, g);
        const double vfz = trilinear_ref(Vz, x, y, z, g);
        double ex = vfy * bz - vfz * by;
        double ey = vfz * bx - vfx * bz;
        double ez = vfx * by - vfy * bx;
        ex = -ex;
        ey = -ey;
        ez = -ez;
        // boris_push, phase_space_gpu_runner.jl:282-323
        const double half = dt / 2.0;
        const double pmx = px + qm * ex * half;
        const double pmy = py + qm * ey * half;
        const double pmz = pz + qm * ez * half;
        const double mp2 = mp * mp, cl2 = cl * cl;
        const double gamma_minus = sqrt(1.0 + (pmx * pmx + pmy * pmy + pmz * pmz) / (mp2 * cl2));
        const double tx = qm * bx * dt / (2.0 * mp * gamma_minus);
        const double ty = qm * by * dt / (2.0 * mp * gamma_minus);
        const double tz = qm * bz * dt / (2.0 * mp * gamma_minus);
        const double t2 = tx * tx + ty * ty + tz * tz;
        const double sx = 2.0 * tx / (1.0 + t2);
        const double sy = 2.0 * ty / (1.0 + t2);
        const double sz = 2.0 * tz / (1.0 + t2);
        const double c1x = pmy * tz - pmz * ty;
        const double c1y = pmz * tx - pmx * tz;
        const double c1z = pmx * ty - pmy * tx;
        const double ppx = pmx + c1x, ppy = pmy + c1y, ppz = pmz + c1z;
        const double c2x = ppy * sz - ppz * sy;
        const double c2y = ppz * sx - ppx * sz;
        const double c2z = ppx * sy - ppy * sx;
        const double ppx2 = pmx + c2x, ppy2 = pmy + c2y, ppz2 = pmz + c2z;
        const double pxn = ppx2 + qm * ex * half;
        const double pyn = ppy2 + qm * ey * half;
        const double pzn = ppz2 + qm * ez * half;
        const double gamma_new = sqrt(1.0 + (pxn * pxn + pyn * pyn + pzn * pzn) / (mp2 * cl2));
        const double vxn = pxn / (gamma_new * mp);
        const double vyn = pyn / (gamma_new * mp);
        const double vzn = pzn / (gamma_new * mp);
        double xn = x + vxn * dt, yn = y + vyn * dt, zn = z + vzn * dt;
        if (periodic) {
            xn = wrap_ref(xn, g.lo[0], g.hi[0]);
            yn = wrap_ref(yn, g.lo[1], g.hi[1]);
            zn = wrap_ref(zn, g.lo[2], g.hi[2]);
        }
        px = pxn;
        py = pyn;
        pz = pzn;
        x = xn;
        y = yn;
        z = zn;
        if (!periodic && (xn < g.lo[0] || xn > g.hi[0] || yn < g.lo[1] || yn > g.hi[1] ||
                          zn < g.lo[2] || zn > g.hi[2])) {
            alive = false;
            atomicAdd(&deaths[si - first_step], 1ull);
        }
    }
    s.x[0][p] = x;
    s.x[1][p] = y;
    s.x[2][p] = z;
    s.p[0][p] = px;
    s.p[1][p] = py;
    s.p[2][p] = pz;
    alive_arr[p] = alive ? 1 : 0;
}

}  // namespace

int64_t launch_push_ref_f64(pt_ens *e, const pt_field *f, const PushArgs &a) {
    GeomD g;
    for (int c = 0; c < 3; ++c) {
        g.lo[c] = f->g.lo[c];
        g.hi[c] = f->g.hi[c];
        g.d[c] = f->g.d[c];
        g.n[c] = f->g.n[c];
    }
    const int K = a.steps_per_launch > 0 ? a.steps_per_launch : 64;
    const unsigned blocks = static_cast<unsigned>((e->n + 255) / 256);
    int64_t launches = 0;
    for (int64_t done = 0; done < a.n_steps; done += K) {
        const int k = static_cast<int>(std::min<int64_t>(K, a.n_steps - done));
        k_push_ref_f64<<<blocks, 256, 0, e->ctx->stream>>>(
            ref_state(e), e->alive, static_cast<const double *>(f->data), g, e->n, a.first_step + done, k,
            a.total_steps, a.stride, a.save_base, a.out_kind, a.out_f64, a.out_a, a.out_b, a.dt, a.q, a.m,
            e->periodic, a.death_base, reinterpret_cast<unsigned long long *>(e->deaths));
        PT_HIP_CHECK(hipGetLastError());
        ++launches;
    }
    return launches;
}

}  // namespace pt
