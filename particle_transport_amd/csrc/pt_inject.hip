// Device injection: init_particles (phase_space_gpu_runner.jl:196-245) with
// one thread per particle.
//
// The reference draws every particle's numbers from one MersenneTwister(seed)
// stream, particle after particle.  A sequential stream cannot be split over
// threads, so each particle here draws from its own counters of the
// counter-based generator Philox4x64-10 (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011): key (seed, 0),
// counter (global particle index, block, 0, 0), four 64-bit words per block.
// Draw j of a particle is word j % 4 of block j / 4, mapped to [0, 1) as
// (word >> 11) * 2^-53.  The per-particle draw order is the reference's.
//
// This file is compiled with -ffp-contract=off: positions follow
// rand * (xmax - xmin) + xmin and the fixed-mu basis follows
// fixed_mu_unit_vector (:170-194) operation by operation, without fused
// multiply-adds, so positions are bit-identical to a plain FP64 evaluation.
#include "pt_internal.hpp"

namespace pt {
namespace {

__device__ __forceinline__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
    constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
        const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
        const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__device__ __forceinline__ double unit_double(uint64_t w) {
    return static_cast<double>(w >> 11) * (1.0 / 9007199254740992.0);
}

struct InjectGeom {
    double lo[3], hi[3], d[3];
    int64_t n[3];
    int64_t nb[3];
    float scale[3];             // 16-bit bricks: decoded = (stored + coarse part) * scale
    const float *coarse;        // 16-bit bricks: coarse-part coefficients
    int32_t cn[3];
};

// Value of B component comp (0..2) at 0-based node (i, j, k) as the field
// stores it, widened to FP64.
template <int STORE>
__device__ __forceinline__ double node_b(const void *__restrict__ data, const InjectGeom &g, int comp,
                                         int64_t i, int64_t j, int64_t k) {
    if constexpr (STORE == PT_STORE_F64_PLANAR || STORE == PT_STORE_F32_PLANAR) {
        const int64_t nn = g.n[0] * g.n[1] * g.n[2];
        const int64_t idx = comp * nn + i + g.n[0] * (j + g.n[1] * k);
        if constexpr (STORE == PT_STORE_F64_PLANAR)
            return static_cast<const double *>(data)[idx];
        else
            return static_cast<double>(static_cast<const float *>(data)[idx]);
    } else {
        const int64_t rec = brick_node_index(g.nb, i, j, k) * 8 + comp;
        if constexpr (STORE == PT_STORE_F32_BRICK) {
            return static_cast<double>(static_cast<const float *>(data)[rec]);
        } else {
            const uint16_t h = static_cast<const uint16_t *>(data)[rec];
            const float v = STORE == PT_STORE_BF16_BRICK ? bf16_bits_to_float(h) : f16_bits_to_float(h);
            return (static_cast<double>(v) + coarse_at_node(g.coarse, g.cn, comp, i, j, k)) *
                   static_cast<double>(g.scale[comp]);
        }
    }
}

__device__ __forceinline__ int64_t clamp_floor(double f, int64_t lo, int64_t hi) {
    const double fl = floor(f);
    if (fl < static_cast<double>(lo)) return lo;
    if (fl > static_cast<double>(hi)) return hi;
    return static_cast<int64_t>(fl);
}

// trilinear_sample (phase_space_gpu_runner.jl:247-276) over the stored nodes,
// in the reference's operation order.
template <int STORE>
__device__ double trilinear_b(const void *__restrict__ data, const InjectGeom &g, int comp, double x,
                              double y, double z) {
    const double fx = (x - g.lo[0]) / g.d[0] + 1.0;
    const double fy = (y - g.lo[1]) / g.d[1] + 1.0;
    const double fz = (z - g.lo[2]) / g.d[2] + 1.0;
    const int64_t i0 = clamp_floor(fx, 1, g.n[0] - 1);
    const int64_t j0 = clamp_floor(fy, 1, g.n[1] - 1);
    const int64_t k0 = clamp_floor(fz, 1, g.n[2] - 1);
    const double tx = fx - static_cast<double>(i0);
    const double ty = fy - static_cast<double>(j0);
    const double tz = fz - static_cast<double>(k0);
    const int64_t i = i0 - 1, j = j0 - 1, k = k0 - 1;
    const double c000 = node_b<STORE>(data, g, comp, i, j, k);
    const double c100 = node_b<STORE>(data, g, comp, i + 1, j, k);
    const double c010 = node_b<STORE>(data, g, comp, i, j + 1, k);
    const double c110 = node_b<STORE>(data, g, comp, i + 1, j + 1, k);
    const double c001 = node_b<STORE>(data, g, comp, i, j, k + 1);
    const double c101 = node_b<STORE>(data, g, comp, i + 1, j, k + 1);
    const double c011 = node_b<STORE>(data, g, comp, i, j + 1, k + 1);
    const double c111 = node_b<STORE>(data, g, comp, i + 1, j + 1, k + 1);
    const double c00 = c000 * (1 - tx) + c100 * tx;
    const double c10 = c010 * (1 - tx) + c110 * tx;
    const double c01 = c001 * (1 - tx) + c101 * tx;
    const double c11 = c011 * (1 - tx) + c111 * tx;
    const double c0 = c00 * (1 - ty) + c10 * ty;
    const double c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

// One thread per particle; out = [x1 x2 x3 p1 p2 p3][count].  bad is set
// where a fixed-mu particle sits at zero or non-finite |B| (its momentum is
// left zero and the host reports the reference's error).
template <int STORE>
__global__ void __launch_bounds__(256)
k_inject(const void *__restrict__ data, InjectGeom g, pt_inject_params ip, int64_t first, int64_t count,
         double *__restrict__ out, int *__restrict__ bad) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const uint64_t i = static_cast<uint64_t>(first + q);
    const int ndraw = (ip.fixed_position ? 0 : 3) + (ip.fixed_mu ? 1 : 2);
    uint64_t w[8];
    w[0] = i;
    w[1] = 0;
    w[2] = 0;
    w[3] = 0;
    philox4x64_10(w, ip.seed, 0);
    if (ndraw > 4) {
        w[4] = i;
        w[5] = 1;
        w[6] = 0;
        w[7] = 0;
        philox4x64_10(w + 4, ip.seed, 0);
    }
    int j = 0;
    double x[3];
    for (int c = 0; c < 3; ++c) x[c] = ip.fixed_position ? ip.position_m[c]
                                                         : unit_double(w[j++]) * (g.hi[c] - g.lo[c]) + g.lo[c];
    const double two_pi = 6.283185307179586;  // 2 * pi rounded to FP64, as Julia's 2 * pi
    double u[3];
    if (ip.fixed_mu) {
        const double bx = trilinear_b<STORE>(data, g, 0, x[0], x[1], x[2]);
        const double by = trilinear_b<STORE>(data, g, 1, x[0], x[1], x[2]);
        const double bz = trilinear_b<STORE>(data, g, 2, x[0], x[1], x[2]);
        const double phi = two_pi * unit_double(w[j++]);
        const double bn = sqrt(bx * bx + by * by + bz * bz);
        if (!isfinite(bn) || bn == 0.0) {
            atomicExch(bad, 1);
            u[0] = u[1] = u[2] = 0.0;
        } else {
            const double bhx = bx / bn, bhy = by / bn, bhz = bz / bn;
            double e1x, e1y, e1z;
            if (fabs(bhx) < 0.9) {
                e1x = 0.0;
                e1y = bhz;
                e1z = -bhy;
            } else {
                e1x = -bhz;
                e1y = 0.0;
                e1z = bhx;
            }
            const double e1n = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
            e1x = e1x / e1n;
            e1y = e1y / e1n;
            e1z = e1z / e1n;
            const double e2x = bhy * e1z - bhz * e1y;
            const double e2y = bhz * e1x - bhx * e1z;
            const double e2z = bhx * e1y - bhy * e1x;
            const double mu0 = ip.mu0;
            const double perp = sqrt(fmax(0.0, 1.0 - mu0 * mu0));
            const double cphi = cos(phi), sphi = sin(phi);
            u[0] = mu0 * bhx + perp * (cphi * e1x + sphi * e2x);
            u[1] = mu0 * bhy + perp * (cphi * e1y + sphi * e2y);
            u[2] = mu0 * bhz + perp * (cphi * e1z + sphi * e2z);
        }
    } else {
        const double z = 2 * unit_double(w[j++]) - 1;
        const double phi = two_pi * unit_double(w[j++]);
        const double r = sqrt(fmax(0.0, 1.0 - z * z));
        u[0] = r * cos(phi);
        u[1] = r * sin(phi);
        u[2] = z;
    }
    for (int c = 0; c < 3; ++c) {
        out[c * count + q] = x[c];
        out[(3 + c) * count + q] = ip.p_magnitude * u[c];
    }
}

}  // namespace

void launch_inject(const pt_field *f, const pt_inject_params &ip, int64_t first, int64_t count, double *out,
                   int *bad, hipStream_t s) {
    InjectGeom g{};
    for (int c = 0; c < 3; ++c) {
        g.lo[c] = f->g.lo[c];
        g.hi[c] = f->g.hi[c];
        g.d[c] = f->g.d[c];
        g.n[c] = f->g.n[c];
        g.nb[c] = f->g.nb[c];
        g.scale[c] = f->sc.scale[c];
        g.cn[c] = f->cn[c];
    }
    g.coarse = f->coarse;
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    switch (f->storage) {
        case PT_STORE_F64_PLANAR:
            k_inject<PT_STORE_F64_PLANAR><<<blocks, 256, 0, s>>>(f->data, g, ip, first, count, out, bad);
            break;
        case PT_STORE_F32_PLANAR:
            k_inject<PT_STORE_F32_PLANAR><<<blocks, 256, 0, s>>>(f->data, g, ip, first, count, out, bad);
            break;
        case PT_STORE_F32_BRICK:
            k_inject<PT_STORE_F32_BRICK><<<blocks, 256, 0, s>>>(f->data, g, ip, first, count, out, bad);
            break;
        case PT_STORE_BF16_BRICK:
            k_inject<PT_STORE_BF16_BRICK><<<blocks, 256, 0, s>>>(f->data, g, ip, first, count, out, bad);
            break;
        case PT_STORE_F16_BRICK:
            k_inject<PT_STORE_F16_BRICK><<<blocks, 256, 0, s>>>(f->data, g, ip, first, count, out, bad);
            break;
        default:
            throw Error{PT_ERR_INVALID, "unknown field storage"};
    }
    PT_HIP_CHECK(hipGetLastError());
}

}  // namespace pt
