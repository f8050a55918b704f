// Host runtime of libpt_hip.so: contexts, fields, ensembles and the
// pt_ens_advance driver, plus the set-up kernels they need (field packing,
// FP64 mean |B|, split-state conversion).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "pt_internal.hpp"

namespace pt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

template <class F>
static int guarded(F &&fn) {
    try {
        fn();
        return PT_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return PT_ERR_MEMORY;
    } catch (...) {
        set_error("unknown C++ exception");
        return PT_ERR_STATE;
    }
}

void *dev_alloc(pt_ctx *ctx, size_t bytes) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{PT_ERR_MEMORY, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " +
                                       hipGetErrorString(e)};
    }
    ctx->bytes_in_use += static_cast<int64_t>(bytes);
    return p;
}
void dev_free(pt_ctx *ctx, void *p, size_t bytes) {
    if (p) {
        (void)hipFree(p);
        ctx->bytes_in_use -= static_cast<int64_t>(bytes);
    }
}

// Size classes of 2^k bytes (at least 4 KiB); free buffers by class.
struct BufferPool {
    std::unordered_multimap<size_t, void *> free;
};

static size_t size_class(size_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

void *pool_get(pt_ctx *ctx, size_t bytes, size_t *granted) {
    if (!ctx->pool) ctx->pool = new BufferPool();
    auto *pool = static_cast<BufferPool *>(ctx->pool);
    const size_t c = size_class(bytes);
    *granted = c;
    auto it = pool->free.find(c);
    if (it != pool->free.end()) {
        void *p = it->second;
        pool->free.erase(it);
        return p;
    }
    return dev_alloc(ctx, c);
}

// Buffers above this size go back to the device at once: their allocation
// cost is small next to the work they serve, and caching them would hold HBM.
constexpr size_t kPoolMaxClass = size_t{64} << 20;

void pool_put(pt_ctx *ctx, void *p, size_t granted) {
    if (!p) return;
    (void)hipStreamSynchronize(ctx->stream);  // idle unless an error cut the call short
    if (granted > kPoolMaxClass) {
        dev_free(ctx, p, granted);
        return;
    }
    static_cast<BufferPool *>(ctx->pool)->free.emplace(granted, p);
}

static void pool_release(pt_ctx *ctx) {
    auto *pool = static_cast<BufferPool *>(ctx->pool);
    if (!pool) return;
    for (auto &kv : pool->free) dev_free(ctx, kv.second, kv.first);
    delete pool;
    ctx->pool = nullptr;
}

void *ctx_scratch(pt_ctx *ctx, size_t bytes) {
    if (bytes > ctx->scratch_bytes) {
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dev_free(ctx, ctx->scratch, ctx->scratch_bytes);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        size_t grow = std::max(bytes, static_cast<size_t>(1) << 20);
        ctx->scratch = dev_alloc(ctx, grow);
        ctx->scratch_bytes = grow;
    }
    return ctx->scratch;
}

// ------------------------------------------------------------- kernels -----

// Planar copy/convert of one component into the field's storage array.
template <class Src, class Dst>
__global__ void k_copy_convert(const Src *__restrict__ src, Dst *__restrict__ dst, int64_t n) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = static_cast<Dst>(src[i]);
}

// Brick packing of component c: node (i,j,k) of the Julia (nx,ny,nz) array
// goes to record brick_node_index(i,j,k), slot c.  rec_words = 8 elements.
template <class Src>
__global__ void k_pack_brick_f32(const Src *__restrict__ src, float *__restrict__ dst, int c,
                                 int64_t nx, int64_t ny, int64_t nz, int64_t nbx, int64_t nby,
                                 int64_t nbz) {
    int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t n = nx * ny * nz;
    if (idx >= n) return;
    int64_t i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
    int64_t nb[3] = {nbx, nby, nbz};
    dst[brick_node_index(nb, i, j, k) * 8 + c] = static_cast<float>(src[idx]);
}

// 16-bit residual of component c: value / scale - coarse part (the coarse
// coefficients are stored in units of scale), rounded to nearest even.
template <class Src, bool BF16>
__global__ void k_pack_brick_16(const Src *__restrict__ src, uint16_t *__restrict__ dst, int c,
                                int64_t nx, int64_t ny, int64_t nz, int64_t nbx, int64_t nby,
                                int64_t nbz, float inv_scale, const float *__restrict__ coarse, int32_t cn0,
                                int32_t cn1, int32_t cn2) {
    int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t n = nx * ny * nz;
    if (idx >= n) return;
    int64_t i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
    int64_t nb[3] = {nbx, nby, nbz};
    const int32_t cn[3] = {cn0, cn1, cn2};
    const double r = static_cast<double>(src[idx]) * static_cast<double>(inv_scale) - coarse_at_node(coarse, cn, c, i, j, k);
    float v = static_cast<float>(r);
    uint16_t bits;
    if (BF16) {
        bits = __bfloat16_as_ushort(__float2bfloat16(v));  // round to nearest even
    } else {
        bits = __half_as_ushort(__float2half_rn(v));
    }
    dst[brick_node_index(nb, i, j, k) * 8 + c] = bits;
}

// Block-level FP64 sum of |B| with a fixed tree; per-block partials are then
// summed in block order by k_sum_partials, so the result is deterministic.
template <class Src>
__global__ void k_abs_b_partials(const Src *__restrict__ bx, const Src *__restrict__ by,
                                 const Src *__restrict__ bz, int64_t n, double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double x = bx[i], y = by[i], z = bz[i];
        acc += sqrt(x * x + y * y + z * z);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double *__restrict__ part, int nparts, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nparts; ++i) s += part[i];
        out[0] = s;
    }
}

// Coarse-part coefficients of component c (16-bit bricks): one thread per
// coarse cell reads the field at the cell's 8 coarse nodes (every 4th node,
// the last cell ending at the grid's last node), forms the trilinear
// polynomial in fine-node units and stores it in FP32.
template <class Src>
__global__ void k_coarse_coeffs(const Src *__restrict__ src, int c, int64_t nx, int64_t ny, int64_t nz,
                                int32_t cn0, int32_t cn1, int32_t cn2, float *__restrict__ table) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncell = static_cast<int64_t>(cn0) * cn1 * cn2;
    if (t >= ncell) return;
    const int64_t I = t % cn0, J = (t / cn0) % cn1, K = t / (static_cast<int64_t>(cn0) * cn1);
    auto node = [](int64_t q, int64_t n) { return (q << kCoarseShift) < n - 1 ? (q << kCoarseShift) : n - 1; };
    const int64_t i0 = node(I, nx), i1 = node(I + 1, nx);
    const int64_t j0 = node(J, ny), j1 = node(J + 1, ny);
    const int64_t k0 = node(K, nz), k1 = node(K + 1, nz);
    const double wx = static_cast<double>(i1 - i0), wy = static_cast<double>(j1 - j0),
                 wz = static_cast<double>(k1 - k0);
    double v[8];
    for (int ci = 0; ci < 8; ++ci) {
        const int64_t i = (ci & 1) ? i1 : i0, j = (ci & 2) ? j1 : j0, k = (ci & 4) ? k1 : k0;
        v[ci] = static_cast<double>(static_cast<float>(src[i + nx * (j + ny * k)]));
    }
    const double dx00 = v[1] - v[0], dx10 = v[3] - v[2], dx01 = v[5] - v[4], dx11 = v[7] - v[6];
    const double dy0 = v[2] - v[0], dy1 = v[6] - v[4];
    float *a = table + (t * 6 + c) * 8;
    a[0] = static_cast<float>(v[0]);
    a[1] = static_cast<float>(dx00 / wx);
    a[2] = static_cast<float>(dy0 / wy);
    a[3] = static_cast<float>((v[4] - v[0]) / wz);
    a[4] = static_cast<float>((dx10 - dx00) / (wx * wy));
    a[5] = static_cast<float>((dx01 - dx00) / (wx * wz));
    a[6] = static_cast<float>((dy1 - dy0) / (wy * wz));
    a[7] = static_cast<float>(((dx11 - dx01) - (dx10 - dx00)) / (wx * wy * wz));
}

// Coarse coefficients of component c times inv_scale (a power of two: exact).
__global__ void k_scale_coarse(float *__restrict__ table, int c, int64_t ncell, float inv_scale) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    float *a = table + (t * 6 + c) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] *= inv_scale;
}

// Largest |value - coarse part| of component c, as the bit pattern of a
// non-negative float (atomic max on the bits orders it as the values).
template <class Src>
__global__ void k_resid_absmax(const Src *__restrict__ src, int c, int64_t nx, int64_t ny, int64_t nz,
                               const float *__restrict__ table, int32_t cn0, int32_t cn1, int32_t cn2,
                               unsigned *__restrict__ out_bits) {
    __shared__ float red[256];
    const int32_t cn[3] = {cn0, cn1, cn2};
    const int64_t n = nx * ny * nz;
    float m = 0.0f;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
         idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
        const double r = static_cast<double>(src[idx]) - coarse_at_node(table, cn, c, i, j, k);
        m = fmaxf(m, static_cast<float>(fabs(r)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out_bits, __float_as_uint(red[0]));
}

// SI state -> split state (split_coord per axis).
__global__ void k_split_from_si(const double *__restrict__ x, const double *__restrict__ p,
                                int64_t n, SplitState s, double d0, double d1, double d2,
                                int32_t nc0, int32_t nc1, int32_t nc2, int periodic, double pnorm) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d[3] = {d0, d1, d2};
    const int32_t nc[3] = {nc0, nc1, nc2};
    double p2 = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double pc = p[c * n + i] / pnorm;
        p2 += pc * pc;
    }
    s.P2[i] = p2;
    for (int c = 0; c < 3; ++c) {
        int32_t cell;
        float off, comp;
        split_coord(x[c * n + i] / d[c], nc[c], periodic != 0, cell, off, comp);
        s.cell[c][i] = cell;
        s.off[c][i] = off;
        s.comp[c][i] = comp;
        s.P[c][i] = static_cast<float>(p[c * n + i] / pnorm);
    }
}

__global__ void k_split_to_si(SplitState s, const int32_t *__restrict__ perm, const uint8_t *__restrict__ alive,
                              uint8_t *__restrict__ alive_out, int64_t n, double *__restrict__ x,
                              double *__restrict__ p, double d0, double d1, double d2, double pnorm) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = perm ? perm[i] : i;  // original particle index of slot i
    double d[3] = {d0, d1, d2};
    // the pushers rescale |U| on a cadence; the readout applies the exact
    // magnitude sqrt(P2) carried in FP64
    double u2 = 0.0;
    for (int c = 0; c < 3; ++c) u2 += static_cast<double>(s.P[c][i]) * static_cast<double>(s.P[c][i]);
    const double sc = u2 > 0.0 ? sqrt(s.P2[i] / u2) : 1.0;
    for (int c = 0; c < 3; ++c) {
        x[c * n + o] = (static_cast<double>(s.cell[c][i]) + static_cast<double>(s.off[c][i]) +
                        static_cast<double>(s.comp[c][i])) * d[c];
        p[c * n + o] = static_cast<double>(s.P[c][i]) * sc * pnorm;
    }
    alive_out[o] = alive[i];
}

static inline unsigned grid_for(int64_t n, int block = 256) {
    return static_cast<unsigned>((n + block - 1) / block);
}

void launch_pack_field(pt_field *f, const void *const src[6], int src_f64, hipStream_t s) {
    const int64_t nx = f->g.n[0], ny = f->g.n[1], nz = f->g.n[2];
    const int64_t n = nx * ny * nz;
    for (int c = 0; c < 6; ++c) {
        if (!src[c]) continue;  // components not present in this pass
        switch (f->storage) {
            case PT_STORE_F64_PLANAR: {
                double *dst = static_cast<double *>(f->data) + c * n;
                if (src_f64)
                    k_copy_convert<<<grid_for(n), 256, 0, s>>>(static_cast<const double *>(src[c]), dst, n);
                else
                    k_copy_convert<<<grid_for(n), 256, 0, s>>>(static_cast<const float *>(src[c]), dst, n);
                break;
            }
            case PT_STORE_F32_PLANAR: {
                float *dst = static_cast<float *>(f->data) + c * n;
                if (src_f64)
                    k_copy_convert<<<grid_for(n), 256, 0, s>>>(static_cast<const double *>(src[c]), dst, n);
                else
                    k_copy_convert<<<grid_for(n), 256, 0, s>>>(static_cast<const float *>(src[c]), dst, n);
                break;
            }
            case PT_STORE_F32_BRICK: {
                float *dst = static_cast<float *>(f->data);
                if (src_f64)
                    k_pack_brick_f32<<<grid_for(n), 256, 0, s>>>(static_cast<const double *>(src[c]), dst, c,
                                                                nx, ny, nz, f->g.nb[0], f->g.nb[1], f->g.nb[2]);
                else
                    k_pack_brick_f32<<<grid_for(n), 256, 0, s>>>(static_cast<const float *>(src[c]), dst, c,
                                                                nx, ny, nz, f->g.nb[0], f->g.nb[1], f->g.nb[2]);
                break;
            }
            case PT_STORE_BF16_BRICK:
            case PT_STORE_F16_BRICK: {
                uint16_t *dst = static_cast<uint16_t *>(f->data);
                const float inv = 1.0f / f->sc.scale[c];
                const bool bf = f->storage == PT_STORE_BF16_BRICK;
                auto pack = [&](auto *sp) {
                    using Src = std::remove_cv_t<std::remove_pointer_t<decltype(sp)>>;
                    if (bf)
                        k_pack_brick_16<Src, true><<<grid_for(n), 256, 0, s>>>(
                            sp, dst, c, nx, ny, nz, f->g.nb[0], f->g.nb[1], f->g.nb[2], inv, f->coarse, f->cn[0],
                            f->cn[1], f->cn[2]);
                    else
                        k_pack_brick_16<Src, false><<<grid_for(n), 256, 0, s>>>(
                            sp, dst, c, nx, ny, nz, f->g.nb[0], f->g.nb[1], f->g.nb[2], inv, f->coarse, f->cn[0],
                            f->cn[1], f->cn[2]);
                };
                if (src_f64)
                    pack(static_cast<const double *>(src[c]));
                else
                    pack(static_cast<const float *>(src[c]));
                break;
            }
        }
        PT_HIP_CHECK(hipGetLastError());
    }
}

void launch_mean_abs_b(pt_ctx *ctx, int64_t n, int src_f64, const void *bx, const void *by,
                       const void *bz, double *out_host) {
    const int nparts = 1024;
    double *part = static_cast<double *>(ctx_scratch(ctx, (nparts + 1) * sizeof(double)));
    if (src_f64)
        k_abs_b_partials<<<nparts, 256, 0, ctx->stream>>>(static_cast<const double *>(bx),
                                                          static_cast<const double *>(by),
                                                          static_cast<const double *>(bz), n, part);
    else
        k_abs_b_partials<<<nparts, 256, 0, ctx->stream>>>(static_cast<const float *>(bx),
                                                          static_cast<const float *>(by),
                                                          static_cast<const float *>(bz), n, part);
    PT_HIP_CHECK(hipGetLastError());
    k_sum_partials<<<1, 64, 0, ctx->stream>>>(part, nparts, part + nparts);
    PT_HIP_CHECK(hipGetLastError());
    double sum = 0.0;
    PT_HIP_CHECK(hipMemcpyAsync(&sum, part + nparts, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out_host = sum / static_cast<double>(n);
}

// Coarse part of component c and the scale of its 16-bit residual (the
// largest |value - coarse part|; 1 where the field equals its coarse part).
static void build_coarse_component(pt_field *f, int c, int src_f64, const void *src) {
    pt_ctx *ctx = f->ctx;
    const int64_t nx = f->g.n[0], ny = f->g.n[1], nz = f->g.n[2];
    const int64_t ncell = static_cast<int64_t>(f->cn[0]) * f->cn[1] * f->cn[2];
    unsigned *bits = static_cast<unsigned *>(ctx_scratch(ctx, 256));
    PT_HIP_CHECK(hipMemsetAsync(bits, 0, sizeof(unsigned), ctx->stream));
    auto run = [&](auto *sp) {
        using Src = std::remove_cv_t<std::remove_pointer_t<decltype(sp)>>;
        k_coarse_coeffs<Src><<<grid_for(ncell), 256, 0, ctx->stream>>>(sp, c, nx, ny, nz, f->cn[0], f->cn[1],
                                                                         f->cn[2], f->coarse);
        PT_HIP_CHECK(hipGetLastError());
        k_resid_absmax<Src><<<1024, 256, 0, ctx->stream>>>(sp, c, nx, ny, nz, f->coarse, f->cn[0], f->cn[1],
                                                            f->cn[2], bits);
        PT_HIP_CHECK(hipGetLastError());
    };
    if (src_f64)
        run(static_cast<const double *>(src));
    else
        run(static_cast<const float *>(src));
    unsigned h = 0;
    PT_HIP_CHECK(hipMemcpyAsync(&h, bits, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float m;
    std::memcpy(&m, &h, sizeof(float));
    // A power of two at least the largest residual: the coarse coefficients
    // are then stored in units of it exactly, and the decode multiplies the
    // coarse part and the residual by the same factor.
    int e = 0;
    std::frexp(static_cast<double>(m), &e);
    const double scale = m > 0.0f && std::isfinite(m) ? std::ldexp(1.0, e) : 1.0;
    f->sc.scale[c] = static_cast<float>(scale);
    k_scale_coarse<<<grid_for(ncell), 256, 0, ctx->stream>>>(f->coarse, c, ncell, static_cast<float>(1.0 / scale));
    PT_HIP_CHECK(hipGetLastError());
}

void launch_split_from_si(pt_ens *e, const double *dx, const double *dp) {
    SplitState s = split_state(e);
    const Geom &g = e->field->g;
    k_split_from_si<<<grid_for(e->n), 256, 0, e->ctx->stream>>>(
        dx, dp, e->n, s, g.d[0], g.d[1], g.d[2], static_cast<int32_t>(g.n[0] - 1), static_cast<int32_t>(g.n[1] - 1),
        static_cast<int32_t>(g.n[2] - 1), e->periodic, kMp * kC);
    PT_HIP_CHECK(hipGetLastError());
}

void launch_split_to_si(pt_ens *e, double *dx, double *dp, uint8_t *alive_out) {
    SplitState s = split_state(e);
    const Geom &g = e->field->g;
    k_split_to_si<<<grid_for(e->n), 256, 0, e->ctx->stream>>>(s, e->perm, e->alive, alive_out, e->n, dx, dp, g.d[0],
                                                             g.d[1], g.d[2], kMp * kC);
    PT_HIP_CHECK(hipGetLastError());
}

static size_t storage_bytes(int storage, const Geom &g) {
    const size_t n = static_cast<size_t>(g.n[0] * g.n[1] * g.n[2]);
    const size_t nbrick = static_cast<size_t>(g.nb[0] * g.nb[1] * g.nb[2]) * 512;
    switch (storage) {
        case PT_STORE_F64_PLANAR: return 6 * n * 8;
        case PT_STORE_F32_PLANAR: return 6 * n * 4;
        case PT_STORE_F32_BRICK: return nbrick * 32;
        case PT_STORE_BF16_BRICK:
        case PT_STORE_F16_BRICK: return nbrick * 16;
    }
    throw Error{PT_ERR_INVALID, "unknown field storage " + std::to_string(storage)};
}

static int64_t saves_in_range(int64_t first, int64_t count, int64_t stride) {
    // number of si in [first, first+count) with (si-1) % stride == 0
    if (count <= 0) return 0;
    int64_t a = first - 1, b = first + count - 2;  // 0-based range [a, b]
    int64_t lo = (a + stride - 1) / stride;
    int64_t hi = b / stride;
    return hi >= lo ? hi - lo + 1 : 0;
}

static void check_pusher_field(int pusher, const pt_field *field, const char *who) {
    PT_REQUIRE(pusher == PT_PUSH_REF_F64 || pusher == PT_PUSH_PLAIN_F32 || pusher == PT_PUSH_FAST,
               std::string(who) + ": unknown pusher");
    if (pusher == PT_PUSH_REF_F64)
        PT_REQUIRE(field->storage == PT_STORE_F64_PLANAR, "PT_PUSH_REF_F64 needs a PT_STORE_F64_PLANAR field");
    if (pusher == PT_PUSH_PLAIN_F32)
        PT_REQUIRE(field->storage == PT_STORE_F32_PLANAR, "PT_PUSH_PLAIN_F32 needs a PT_STORE_F32_PLANAR field");
    if (pusher == PT_PUSH_FAST)
        PT_REQUIRE(field->storage != PT_STORE_F64_PLANAR, "PT_PUSH_FAST needs an FP32 or 16-bit field");
}

// The reference's checks on the injection settings (init_particles,
// phase_space_gpu_runner.jl:215-219; fixed_injection_position :164-166).
void check_inject_params(const pt_inject_params &ip, const pt_field *field) {
    PT_REQUIRE(ip.fixed_position == 0 || ip.fixed_position == 1, "pt_inject_params.fixed_position must be 0 or 1");
    PT_REQUIRE(ip.fixed_mu == 0 || ip.fixed_mu == 1, "pt_inject_params.fixed_mu must be 0 or 1");
    PT_REQUIRE(std::isfinite(ip.p_magnitude), "pt_inject_params.p_magnitude must be finite");
    if (ip.fixed_position)
        for (int c = 0; c < 3; ++c)
            PT_REQUIRE(std::isfinite(ip.position_m[c]) && field->g.lo[c] <= ip.position_m[c] &&
                           ip.position_m[c] <= field->g.hi[c],
                       "fixed injection position is outside the simulation box");
    if (ip.fixed_mu) PT_REQUIRE(ip.mu0 >= -1.0 && ip.mu0 <= 1.0, "injection_mu0 must be in [-1, 1]");
}

// Allocates an ensemble and its state.  fill(si, flag) enqueues on the
// context stream the writes of the FP64 SI state [x1 x2 x3 p1 p2 p3][n] into
// the device array si; flag is one int of device scratch it may use.  The
// split pushers' state is converted from si on the device.
template <class Fill>
static pt_ens *make_ensemble(pt_ctx *ctx, const pt_field *field, int pusher, int64_t n, int periodic,
                             Fill &&fill) {
    const size_t N = static_cast<size_t>(n);
    auto *e = new pt_ens();
    e->ctx = ctx;
    e->field = field;
    e->pusher = pusher;
    e->n = n;
    e->periodic = periodic ? 1 : 0;
    e->n_alive = n;
    try {
        e->state_bytes = pusher == PT_PUSH_REF_F64 ? 6 * N * 8 : kSplitStateBytes * N;
        e->state = dev_alloc(ctx, e->state_bytes);
        e->alive = static_cast<uint8_t *>(dev_alloc(ctx, N));
        PT_HIP_CHECK(hipMemsetAsync(e->alive, 1, N, ctx->stream));
        const size_t si_bytes = pusher == PT_PUSH_REF_F64 ? 0 : 6 * N * 8;
        char *scr = static_cast<char *>(ctx_scratch(ctx, si_bytes + 256));
        double *si = pusher == PT_PUSH_REF_F64 ? static_cast<double *>(e->state) : reinterpret_cast<double *>(scr);
        fill(si, reinterpret_cast<int *>(scr + si_bytes));
        if (pusher != PT_PUSH_REF_F64) launch_split_from_si(e, si, si + 3 * N);
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, e->state, e->state_bytes);
        dev_free(ctx, e->alive, N);
        delete e;
        throw;
    }
    return e;
}

}  // namespace pt

using namespace pt;

extern "C" {

int pt_abi_version(void) { return PT_ABI_VERSION; }

const char *pt_last_error(void) { return pt::g_last_error.c_str(); }

int pt_ctx_create(int device, pt_ctx **out) {
    return guarded([&] {
        PT_REQUIRE(out != nullptr, "pt_ctx_create: out is null");
        int count = 0;
        PT_HIP_CHECK(hipGetDeviceCount(&count));
        PT_REQUIRE(device >= 0 && device < count, "pt_ctx_create: device ordinal out of range");
        PT_HIP_CHECK(hipSetDevice(device));
        noalloc_scratch_init(device);
        auto *c = new pt_ctx();
        c->device = device;
        // Default-priority, non-blocking stream owned by the context.
        PT_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        PT_HIP_CHECK(hipEventCreate(&c->ev0));
        PT_HIP_CHECK(hipEventCreate(&c->ev1));
        *out = c;
    });
}

int pt_ctx_destroy(pt_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        pool_release(ctx);
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int pt_ctx_bytes_in_use(const pt_ctx *ctx, int64_t *bytes) {
    return guarded([&] {
        PT_REQUIRE(ctx && bytes, "pt_ctx_bytes_in_use: null argument");
        *bytes = ctx->bytes_in_use;
    });
}

int pt_device_memory(const pt_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes) {
    return guarded([&] {
        PT_REQUIRE(ctx && free_bytes && total_bytes, "pt_device_memory: null argument");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        size_t f = 0, t = 0;
        PT_HIP_CHECK(hipMemGetInfo(&f, &t));
        *free_bytes = static_cast<int64_t>(f);
        *total_bytes = static_cast<int64_t>(t);
    });
}

int pt_buffer_alloc(pt_ctx *ctx, int64_t bytes, void **dev_ptr) {
    return guarded([&] {
        PT_REQUIRE(ctx && dev_ptr, "pt_buffer_alloc: null argument");
        PT_REQUIRE(bytes >= 0, "pt_buffer_alloc: negative size");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        *dev_ptr = dev_alloc(ctx, static_cast<size_t>(bytes));
    });
}

int pt_buffer_free(pt_ctx *ctx, void *dev_ptr, int64_t bytes) {
    return guarded([&] {
        PT_REQUIRE(ctx, "pt_buffer_free: null context");
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dev_free(ctx, dev_ptr, static_cast<size_t>(bytes));
    });
}

int pt_memcpy(pt_ctx *ctx, void *dst, const void *src, int64_t bytes, int direction) {
    return guarded([&] {
        PT_REQUIRE(ctx && dst && src, "pt_memcpy: null argument");
        PT_REQUIRE(bytes >= 0, "pt_memcpy: negative size");
        PT_REQUIRE(direction >= 0 && direction <= 2, "pt_memcpy: direction must be 0, 1 or 2");
        const hipMemcpyKind k = direction == 0 ? hipMemcpyHostToDevice
                                : direction == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        PT_HIP_CHECK(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), k, ctx->stream));
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int pt_memcpy_2d(pt_ctx *ctx, void *dst, int64_t dst_pitch, const void *src, int64_t src_pitch, int64_t width,
                 int64_t rows, int direction) {
    return guarded([&] {
        PT_REQUIRE(ctx && dst && src, "pt_memcpy_2d: null argument");
        PT_REQUIRE(width >= 0 && rows >= 0, "pt_memcpy_2d: negative size");
        PT_REQUIRE(dst_pitch >= width && src_pitch >= width, "pt_memcpy_2d: a pitch is narrower than a row");
        PT_REQUIRE(direction >= 0 && direction <= 2, "pt_memcpy_2d: direction must be 0, 1 or 2");
        if (width == 0 || rows == 0) return;
        const hipMemcpyKind k = direction == 0 ? hipMemcpyHostToDevice
                                : direction == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        PT_HIP_CHECK(hipMemcpy2DAsync(dst, static_cast<size_t>(dst_pitch), src, static_cast<size_t>(src_pitch),
                                      static_cast<size_t>(width), static_cast<size_t>(rows), k, ctx->stream));
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int pt_host_alloc(int64_t bytes, void **host_ptr) {
    return guarded([&] {
        PT_REQUIRE(host_ptr && bytes >= 0, "pt_host_alloc: bad argument");
        PT_HIP_CHECK(hipHostMalloc(host_ptr, static_cast<size_t>(bytes ? bytes : 1), hipHostMallocDefault));
    });
}

int pt_host_free(void *host_ptr) {
    return guarded([&] {
        if (host_ptr) PT_HIP_CHECK(hipHostFree(host_ptr));
    });
}

int pt_field_create(pt_ctx *ctx, int64_t nx, int64_t ny, int64_t nz, double box_length_m,
                    int storage, int src_dtype, int src_on_device, const void *const comps[6],
                    pt_field **out) {
    return guarded([&] {
        PT_REQUIRE(ctx && comps && out, "pt_field_create: null argument");
        PT_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "pt_field_create: need at least 2 nodes per axis");
        PT_REQUIRE(nx < (1LL << 30) && ny < (1LL << 30) && nz < (1LL << 30), "pt_field_create: grid too large");
        PT_REQUIRE(std::isfinite(box_length_m) && box_length_m > 0, "pt_field_create: box length must be finite and positive");
        PT_REQUIRE(src_dtype == PT_F64 || src_dtype == PT_F32, "pt_field_create: src_dtype must be PT_F64 or PT_F32");
        for (int c = 0; c < 6; ++c) PT_REQUIRE(comps[c] != nullptr, "pt_field_create: null component");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        auto *f = new pt_field();
        f->ctx = ctx;
        f->storage = storage;
        int64_t nn[3] = {nx, ny, nz};
        for (int c = 0; c < 3; ++c) {
            f->g.n[c] = nn[c];
            f->g.lo[c] = 0.0;
            f->g.hi[c] = box_length_m;
            f->g.d[c] = box_length_m / static_cast<double>(nn[c] - 1);
            f->g.nb[c] = (nn[c] + 7) / 8;
        }
        f->g.L = box_length_m;
        if (storage == PT_STORE_F32_BRICK || storage == PT_STORE_BF16_BRICK || storage == PT_STORE_F16_BRICK) {
            const int64_t records = f->g.nb[0] * f->g.nb[1] * f->g.nb[2] * 512;
            if (records > INT32_MAX) {
                delete f;
                throw Error{PT_ERR_INVALID, "pt_field_create: brick storage holds at most 2^31 node records "
                                             "(about 1288^3 nodes); this grid needs " + std::to_string(records)};
            }
        }
        void *staging = nullptr;
        size_t staging_bytes = 0;
        try {
            f->bytes = storage_bytes(storage, f->g);
            const int64_t n = nx * ny * nz;
            const size_t esz = src_dtype == PT_F64 ? 8 : 4;
            // Host sources are staged three components at a time in a
            // temporary device buffer of their own (the context scratch is
            // used by the reductions below).
            if (!src_on_device) {
                staging_bytes = 3 * static_cast<size_t>(n) * esz;
                staging = dev_alloc(ctx, staging_bytes);
            }
            f->data = dev_alloc(ctx, f->bytes);
            PT_HIP_CHECK(hipMemsetAsync(f->data, 0, f->bytes, ctx->stream));
            if (storage == PT_STORE_BF16_BRICK || storage == PT_STORE_F16_BRICK) {
                for (int c = 0; c < 3; ++c)
                    f->cn[c] = static_cast<int32_t>((nn[c] - 1 + kCoarseWidth - 1) / kCoarseWidth);
                f->coarse_bytes = static_cast<size_t>(f->cn[0]) * f->cn[1] * f->cn[2] * 48 * sizeof(float);
                f->coarse = static_cast<float *>(dev_alloc(ctx, f->coarse_bytes));
            }
            for (int half = 0; half < 2; ++half) {
                const void *dsrc[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
                for (int c = 0; c < 3; ++c) {
                    const int comp = 3 * half + c;
                    if (src_on_device) {
                        dsrc[comp] = comps[comp];
                    } else {
                        void *dst = static_cast<char *>(staging) + c * n * esz;
                        PT_HIP_CHECK(hipMemcpyAsync(dst, comps[comp], n * esz, hipMemcpyHostToDevice, ctx->stream));
                        dsrc[comp] = dst;
                    }
                }
                if (half == 0)
                    launch_mean_abs_b(ctx, n, src_dtype == PT_F64, dsrc[0], dsrc[1], dsrc[2], &f->B0);
                for (int c = 0; c < 3; ++c) {
                    const int comp = 3 * half + c;
                    if (storage == PT_STORE_BF16_BRICK || storage == PT_STORE_F16_BRICK) {
                        build_coarse_component(f, comp, src_dtype == PT_F64, dsrc[comp]);
                    } else {
                        f->sc.scale[comp] = 1.0f;
                    }
                }
                launch_pack_field(f, dsrc, src_dtype == PT_F64, ctx->stream);
                PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            }
            dev_free(ctx, staging, staging_bytes);
            staging = nullptr;
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            dev_free(ctx, staging, staging_bytes);
            if (f->data) dev_free(ctx, f->data, f->bytes);
            if (f->coarse) dev_free(ctx, f->coarse, f->coarse_bytes);
            delete f;
            throw;
        }
        *out = f;
    });
}

int pt_field_destroy(pt_field *field) {
    return guarded([&] {
        if (!field) return;
        (void)hipStreamSynchronize(field->ctx->stream);
        dev_free(field->ctx, field->data, field->bytes);
        if (field->coarse) dev_free(field->ctx, field->coarse, field->coarse_bytes);
        if (field->cells) dev_free(field->ctx, field->cells, field->cells_bytes);
        delete field;
    });
}

int pt_field_mean_abs_b(const pt_field *field, double *B0_T) {
    return guarded([&] {
        PT_REQUIRE(field && B0_T, "pt_field_mean_abs_b: null argument");
        *B0_T = field->B0;
    });
}

int pt_mean_abs_b(pt_ctx *ctx, int64_t n, int src_dtype, int src_on_device, const void *bx,
                  const void *by, const void *bz, double *B0_T) {
    return guarded([&] {
        PT_REQUIRE(ctx && bx && by && bz && B0_T, "pt_mean_abs_b: null argument");
        PT_REQUIRE(n > 0, "pt_mean_abs_b: n must be positive");
        PT_REQUIRE(src_dtype == PT_F64 || src_dtype == PT_F32, "pt_mean_abs_b: bad dtype");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        const size_t esz = src_dtype == PT_F64 ? 8 : 4;
        if (src_on_device) {
            launch_mean_abs_b(ctx, n, src_dtype == PT_F64, bx, by, bz, B0_T);
            return;
        }
        // Stage after the reduction's partials (scratch head is used by the reduction).
        const size_t head = 4096 * sizeof(double);
        char *st = static_cast<char *>(ctx_scratch(ctx, head + 3 * n * esz)) + head;
        PT_HIP_CHECK(hipMemcpyAsync(st, bx, n * esz, hipMemcpyHostToDevice, ctx->stream));
        PT_HIP_CHECK(hipMemcpyAsync(st + n * esz, by, n * esz, hipMemcpyHostToDevice, ctx->stream));
        PT_HIP_CHECK(hipMemcpyAsync(st + 2 * n * esz, bz, n * esz, hipMemcpyHostToDevice, ctx->stream));
        launch_mean_abs_b(ctx, n, src_dtype == PT_F64, st, st + n * esz, st + 2 * n * esz, B0_T);
    });
}

int pt_ens_create(pt_ctx *ctx, const pt_field *field, int pusher, int64_t n, int src_dtype,
                  int src_on_device, const void *const x[3], const void *const p[3], int periodic,
                  pt_ens **out) {
    return guarded([&] {
        PT_REQUIRE(ctx && field && x && p && out, "pt_ens_create: null argument");
        PT_REQUIRE(n >= 1 && n < (1LL << 31), "pt_ens_create: particle count out of range");
        PT_REQUIRE(src_dtype == PT_F64 || src_dtype == PT_F32, "pt_ens_create: bad dtype");
        check_pusher_field(pusher, field, "pt_ens_create");
        for (int c = 0; c < 3; ++c) PT_REQUIRE(x[c] && p[c], "pt_ens_create: null state component");
        PT_HIP_CHECK(hipSetDevice(ctx->device));

        // Gather the state as FP64 on the host, check it, stage it on the device.
        const size_t N = static_cast<size_t>(n);
        std::vector<double> h(6 * N);
        const size_t esz = src_dtype == PT_F64 ? 8 : 4;
        std::vector<char> tmp(src_on_device ? N * esz : 0);
        for (int c = 0; c < 6; ++c) {
            const void *srcp = c < 3 ? x[c] : p[c - 3];
            const void *hsrc = srcp;
            if (src_on_device) {
                PT_HIP_CHECK(hipMemcpy(tmp.data(), srcp, N * esz, hipMemcpyDeviceToHost));
                hsrc = tmp.data();
            }
            for (size_t i = 0; i < N; ++i)
                h[c * N + i] = src_dtype == PT_F64 ? static_cast<const double *>(hsrc)[i]
                                                   : static_cast<const float *>(hsrc)[i];
        }
        const Geom &g = field->g;
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < N; ++i) {
                double v = h[c * N + i];
                PT_REQUIRE(std::isfinite(v) && v >= g.lo[c] && v <= g.hi[c],
                           "pt_ens_create: initial position outside the simulation box");
                PT_REQUIRE(std::isfinite(h[(3 + c) * N + i]), "pt_ens_create: non-finite momentum");
            }
        *out = make_ensemble(ctx, field, pusher, n, periodic, [&](double *si, int *) {
            PT_HIP_CHECK(hipMemcpyAsync(si, h.data(), 6 * N * 8, hipMemcpyHostToDevice, ctx->stream));
        });
    });
}

int pt_inject(pt_ctx *ctx, const pt_field *field, int64_t first_index, int64_t n,
              const pt_inject_params *params, int out_on_device, double *out) {
    return guarded([&] {
        PT_REQUIRE(ctx && field && params && out, "pt_inject: null argument");
        PT_REQUIRE(n >= 1 && n < (1LL << 31), "pt_inject: particle count out of range");
        PT_REQUIRE(first_index >= 0 && first_index <= INT64_MAX - n, "pt_inject: first_index out of range");
        check_inject_params(*params, field);
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        const size_t bytes = 6 * static_cast<size_t>(n) * 8;
        char *scr = static_cast<char *>(ctx_scratch(ctx, (out_on_device ? 0 : bytes) + 256));
        double *dst = out_on_device ? out : reinterpret_cast<double *>(scr);
        int *flag = reinterpret_cast<int *>(scr + (out_on_device ? 0 : bytes));
        int bad = 0;
        PT_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
        launch_inject(field, *params, first_index, n, dst, flag, ctx->stream);
        if (!out_on_device) PT_HIP_CHECK(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PT_HIP_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        PT_REQUIRE(!bad, "cannot use fixed-mu injection where local B is zero or non-finite");
    });
}

int pt_ens_create_injected(pt_ctx *ctx, const pt_field *field, int pusher, int64_t n,
                           const pt_inject_params *params, int periodic, pt_ens **out) {
    return guarded([&] {
        PT_REQUIRE(ctx && field && params && out, "pt_ens_create_injected: null argument");
        PT_REQUIRE(n >= 1 && n < (1LL << 31), "pt_ens_create_injected: particle count out of range");
        check_pusher_field(pusher, field, "pt_ens_create_injected");
        check_inject_params(*params, field);
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        int bad = 0;
        pt_ens *e = make_ensemble(ctx, field, pusher, n, periodic, [&](double *si, int *flag) {
            PT_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
            launch_inject(field, *params, 0, n, si, flag, ctx->stream);
            PT_HIP_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        });
        if (bad) {
            (void)pt_ens_destroy(e);
            throw Error{PT_ERR_INVALID, "cannot use fixed-mu injection where local B is zero or non-finite"};
        }
        *out = e;
    });
}

int pt_ens_destroy(pt_ens *ens) {
    return guarded([&] {
        if (!ens) return;
        pt_ctx *ctx = ens->ctx;
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, ens->state, ens->state_bytes);
        dev_free(ctx, ens->alive, static_cast<size_t>(ens->n));
        dev_free(ctx, ens->deaths, static_cast<size_t>(ens->deaths_cap) * 8);
        if (ens->perm) dev_free(ctx, ens->perm, static_cast<size_t>(ens->n) * 4);
        dev_free(ctx, ens->state_alt, ens->state_bytes);
        dev_free(ctx, ens->alive_alt, static_cast<size_t>(ens->n));
        dev_free(ctx, ens->perm_alt, ens->perm_alt ? static_cast<size_t>(ens->n) * 4 : 0);
        dev_free(ctx, ens->sort_ws, ens->sort_ws_bytes);
        delete ens;
    });
}

int pt_ens_advance(pt_ens *ens, const pt_field *field, const pt_advance_params *params,
                   void *out_a, void *out_b, int out_dtype, int out_on_device,
                   int64_t out_capacity_saves, int64_t *alive_counts, int64_t *n_saved) {
    return guarded([&] {
        PT_REQUIRE(ens && field && params, "pt_ens_advance: null argument");
        PT_REQUIRE(field->g.n[0] == ens->field->g.n[0] && field->g.n[1] == ens->field->g.n[1] &&
                       field->g.n[2] == ens->field->g.n[2] && field->g.L == ens->field->g.L,
                   "pt_ens_advance: field geometry differs from the ensemble's");
        if (ens->pusher == PT_PUSH_REF_F64)
            PT_REQUIRE(field->storage == PT_STORE_F64_PLANAR, "PT_PUSH_REF_F64 needs a PT_STORE_F64_PLANAR field");
        if (ens->pusher == PT_PUSH_PLAIN_F32)
            PT_REQUIRE(field->storage == PT_STORE_F32_PLANAR, "PT_PUSH_PLAIN_F32 needs a PT_STORE_F32_PLANAR field");
        if (ens->pusher == PT_PUSH_FAST)
            PT_REQUIRE(field->storage != PT_STORE_F64_PLANAR, "PT_PUSH_FAST needs an FP32 or 16-bit field");
        const pt_advance_params &P = *params;
        PT_REQUIRE(std::isfinite(P.dt_s) && P.dt_s > 0, "pt_ens_advance: dt must be finite and positive");
        PT_REQUIRE(std::isfinite(P.mass_kg) && P.mass_kg > 0, "pt_ens_advance: mass must be positive");
        PT_REQUIRE(std::isfinite(P.charge_C) && P.charge_C != 0, "pt_ens_advance: charge must be non-zero");
        PT_REQUIRE(P.first_step >= 1 && P.n_steps >= 0 && P.total_steps >= 1,
                   "pt_ens_advance: step indices must be >= 1");
        PT_REQUIRE(P.first_step + P.n_steps - 1 <= P.total_steps, "pt_ens_advance: steps past total_steps");
        PT_REQUIRE(P.save_stride >= 1, "trajectory_time_stride must be >= 1");
        PT_REQUIRE(P.out_kind == PT_OUT_NONE || P.out_kind == PT_OUT_PHASE || P.out_kind == PT_OUT_MU,
                   "pt_ens_advance: unknown out_kind");
        PT_REQUIRE(out_dtype == PT_F32 || out_dtype == PT_F64, "pt_ens_advance: out_dtype must be PT_F32 or PT_F64");
        pt_ctx *ctx = ens->ctx;
        PT_HIP_CHECK(hipSetDevice(ctx->device));

        const int64_t nsv = saves_in_range(P.first_step, P.n_steps, P.save_stride);
        const size_t N = static_cast<size_t>(ens->n);
        const size_t esz = out_dtype == PT_F64 ? 8 : 4;
        const size_t comps = P.out_kind == PT_OUT_PHASE ? 3 : 1;
        void *da = nullptr, *db = nullptr;
        char *staging = nullptr;
        if (P.out_kind != PT_OUT_NONE && nsv > 0) {
            PT_REQUIRE(out_a != nullptr && (P.out_kind != PT_OUT_PHASE || out_b != nullptr),
                       "pt_ens_advance: output buffer is null");
            PT_REQUIRE(out_capacity_saves >= nsv, "pt_ens_advance: output buffers hold " +
                                                      std::to_string(out_capacity_saves) + " saves, call writes " +
                                                      std::to_string(nsv));
            if (out_on_device) {
                da = out_a;
                db = out_b;
            } else {
                size_t per = static_cast<size_t>(nsv) * comps * N * esz;
                staging = static_cast<char *>(ctx_scratch(ctx, per * (P.out_kind == PT_OUT_PHASE ? 2 : 1)));
                da = staging;
                db = P.out_kind == PT_OUT_PHASE ? staging + per : nullptr;
            }
        }
        if (ens->deaths_cap < std::max<int64_t>(P.n_steps, 1)) {
            dev_free(ctx, ens->deaths, static_cast<size_t>(ens->deaths_cap) * 8);
            ens->deaths = nullptr;
            ens->deaths_cap = 0;
            int64_t cap = std::max<int64_t>(P.n_steps, 1024);
            ens->deaths = static_cast<int64_t *>(dev_alloc(ctx, static_cast<size_t>(cap) * 8));
            ens->deaths_cap = cap;
        }
        PT_HIP_CHECK(hipMemsetAsync(ens->deaths, 0, static_cast<size_t>(std::max<int64_t>(P.n_steps, 1)) * 8, ctx->stream));

        PushArgs a{};
        a.first_step = P.first_step;
        a.n_steps = P.n_steps;
        a.total_steps = P.total_steps;
        a.stride = P.save_stride;
        a.save_base = (P.first_step - 1 + P.save_stride - 1) / P.save_stride;
        a.out_kind = (nsv > 0) ? P.out_kind : PT_OUT_NONE;
        a.out_f64 = out_dtype == PT_F64;
        a.out_a = da;
        a.out_b = db;
        a.out_cap = nsv;
        a.dt = P.dt_s;
        a.q = P.charge_C;
        a.m = P.mass_kg;
        a.steps_per_launch = P.steps_per_launch;
        a.death_base = P.first_step;
        PT_REQUIRE(P.sort_every_steps >= 0, "pt_ens_advance: sort_every_steps must be >= 0");
        PT_REQUIRE(P.sort_every_steps == 0 || ens->pusher != PT_PUSH_REF_F64,
                   "pt_ens_advance: the FP64 reference pusher does not re-sort");

        // Segments between Morton re-sorts; the timing covers the push launches only.
        const int64_t seg = P.sort_every_steps > 0 ? P.sort_every_steps : std::max<int64_t>(P.n_steps, 1);
        int64_t launches = 0;
        double ms_total = 0.0;
        for (int64_t done = 0; done < P.n_steps; done += seg) {
            if (P.sort_every_steps > 0) morton_sort(ens);
            PushArgs s = a;
            s.first_step = P.first_step + done;
            s.n_steps = std::min(seg, P.n_steps - done);
            PT_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
            switch (ens->pusher) {
                case PT_PUSH_REF_F64: launches += launch_push_ref_f64(ens, field, s); break;
                case PT_PUSH_PLAIN_F32: launches += launch_push_plain_f32(ens, field, s); break;
                case PT_PUSH_FAST: launches += launch_push_fast(ens, field, s); break;
            }
            PT_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
            PT_HIP_CHECK(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            PT_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            ms_total += ms;
        }
        ens->kernel_ms = ms_total;
        ens->launches = launches;

        std::vector<int64_t> deaths(static_cast<size_t>(std::max<int64_t>(P.n_steps, 1)));
        PT_HIP_CHECK(hipMemcpy(deaths.data(), ens->deaths, deaths.size() * 8, hipMemcpyDeviceToHost));
        int64_t alive = ens->n_alive;
        for (int64_t k = 0; k < P.n_steps; ++k) {
            alive -= deaths[static_cast<size_t>(k)];
            if (alive_counts) alive_counts[k] = alive;
        }
        ens->n_alive = alive;

        if (staging) {
            size_t per = static_cast<size_t>(nsv) * comps * N * esz;
            PT_HIP_CHECK(hipMemcpy(out_a, da, per, hipMemcpyDeviceToHost));
            if (P.out_kind == PT_OUT_PHASE) PT_HIP_CHECK(hipMemcpy(out_b, db, per, hipMemcpyDeviceToHost));
        }
        if (n_saved) *n_saved = nsv;
    });
}

int pt_ens_get_state(pt_ens *ens, double *x[3], double *p[3], uint8_t *alive) {
    return guarded([&] {
        PT_REQUIRE(ens, "pt_ens_get_state: null ensemble");
        pt_ctx *ctx = ens->ctx;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        const size_t N = static_cast<size_t>(ens->n);
        std::vector<double> h(6 * N);
        if (ens->pusher == PT_PUSH_REF_F64) {
            PT_HIP_CHECK(hipMemcpyAsync(h.data(), ens->state, 6 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
            if (alive) PT_HIP_CHECK(hipMemcpyAsync(alive, ens->alive, N, hipMemcpyDeviceToHost, ctx->stream));
        } else {
            char *st = static_cast<char *>(ctx_scratch(ctx, 6 * N * 8 + N));
            auto *sd = reinterpret_cast<double *>(st);
            auto *sa = reinterpret_cast<uint8_t *>(st + 6 * N * 8);
            launch_split_to_si(ens, sd, sd + 3 * N, sa);
            PT_HIP_CHECK(hipMemcpyAsync(h.data(), sd, 6 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
            if (alive) PT_HIP_CHECK(hipMemcpyAsync(alive, sa, N, hipMemcpyDeviceToHost, ctx->stream));
        }
        PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 3; ++c) {
            if (x && x[c]) std::memcpy(x[c], h.data() + c * N, N * 8);
            if (p && p[c]) std::memcpy(p[c], h.data() + (3 + c) * N, N * 8);
        }
    });
}

int pt_ens_sort(pt_ens *ens) {
    return guarded([&] {
        PT_REQUIRE(ens, "pt_ens_sort: null ensemble");
        PT_HIP_CHECK(hipSetDevice(ens->ctx->device));
        morton_sort(ens);
    });
}

int pt_ens_last_timing(const pt_ens *ens, double *kernel_ms, int64_t *launches) {
    return guarded([&] {
        PT_REQUIRE(ens, "pt_ens_last_timing: null ensemble");
        if (kernel_ms) *kernel_ms = ens->kernel_ms;
        if (launches) *launches = ens->launches;
    });
}

}  // extern "C"
