// Lag statistics and mu reconstruction on gfx950.
//
//  D_mumu (sliding and injection starts) and D_pp replace the host-threaded
//  accumulators of compute_delta_mu2_dmumu_full.jl:457-654 and
//  compute_dpp_full.jl:247-288.  One lane walks one particle's time series;
//  a workgroup covers 256 particles x G lags.  The mu-bin of the pair start
//  changes rarely along a trajectory, so each lane keeps a run accumulator
//  (count, sum dmu, sum dmu^2) per lag for its current bin in registers,
//  moves it to a one-run stash when the bin changes, and the wave flushes
//  its lanes' stashes into LDS-privatised bins (ds_add) together, when one
//  of them is needed again.
//  Per-particle means are merged across particles with Chan's formula in a
//  fixed tree (wavefront shuffles, then LDS), per-workgroup partials are
//  merged in workgroup order, so the Welford statistics are deterministic.
//  The per-bin sums are accumulated as 128-bit fixed-point integers (2^-80
//  units), whose addition is exact and associative, so they are the same
//  whatever order the lanes and workgroups flush in.  A flush converts a run
//  sum with integer operations on the double's bits and adds it to the LDS
//  bins as three carry-free words with non-returning atomics; each workgroup
//  folds its words into the global 128-bit bins once, at its end.  Integer
//  counts are exact; the remaining sums are FP64.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "pt_internal.hpp"

namespace pt {
namespace {

constexpr int kTPB = 256;
// Lags per workgroup of the D_pp walk: eight over float32 series, four over
// FP64 series, whose samples and sums take twice the registers (eight FP64
// lags held the walk to three waves per SIMD; four give five).
template <class T>
constexpr int kDppLags = sizeof(T) == 8 ? 4 : 8;
constexpr int kGmu = 4;        // lags per workgroup (D_mumu: its run accumulators take more registers)
constexpr int kMaxBins = 256;  // LDS budget: kGmu * bins * 56 B (57 KB at 256 bins)

struct Welford {
    double n, mean, m2;
};

__device__ __forceinline__ Welford chan(Welford a, Welford b) {
    if (b.n == 0.0) return a;
    if (a.n == 0.0) return b;
    const double n = a.n + b.n;
    const double delta = b.mean - a.mean;
    Welford r;
    r.n = n;
    r.mean = a.mean + delta * b.n / n;
    r.m2 = a.m2 + b.m2 + delta * delta * a.n * b.n / n;
    return r;
}

// Block-wide ordered reduction helpers (256 threads = 4 wavefronts of 64).
__device__ __forceinline__ Welford block_welford(Welford w, Welford *sh) {
    for (int o = 32; o > 0; o >>= 1) {
        Welford other;
        other.n = __shfl_down(w.n, o, 64);
        other.mean = __shfl_down(w.mean, o, 64);
        other.m2 = __shfl_down(w.m2, o, 64);
        w = chan(w, other);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = w;
    __syncthreads();
    Welford r{0, 0, 0};
    if (threadIdx.x == 0)
        for (int i = 0; i < kTPB / 64; ++i) r = chan(r, sh[i]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kTPB / 64; ++i) r += sh[i];
    __syncthreads();
    return r;
}

__device__ __forceinline__ long long block_sum_i(long long v, long long *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    long long r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kTPB / 64; ++i) r += sh[i];
    __syncthreads();
    return r;
}

// Linear block id -> (particle tile, lag group).  Hardware dispatches block
// b to XCD b % 8; the lag groups of one tile are given consecutive ids on the
// same XCD, so they walk the same series together out of that XCD's L2
// instead of each reading it from HBM.  tile >= nbx: an idle block.
struct TileGroup {
    int tile, group;
};
__device__ __forceinline__ TileGroup tile_group(int ngy) {
    const int b = static_cast<int>(blockIdx.x);
    const int r = b >> 3;
    return {(r / ngy) * 8 + (b & 7), r % ngy};
}
__host__ inline unsigned tile_group_blocks(int nbx, int ngy) {
    return static_cast<unsigned>(((nbx + 7) / 8) * 8 * ngy);
}

// ---------------------------------------------------------------- D_mumu ---
//
// Bins: bin b (0-based) holds float v with lo[b] <= v < lo[b+1] (hi_last for
// the last bin).  lo[b] is the smallest float >= edge b, and hi_last the
// successor of the largest float <= the last edge, so for every float mu these
// comparisons equal mu_bin_index_full's FP64 comparisons
// (compute_delta_mu2_dmumu_full.jl:308-318), right-closed last bin included.
// 128-bit two's-complement fixed point in units of 2^-80: {lo, hi} words.
// A bin's sum over a whole call stays below 2^47 in magnitude (|dmu| <= 2,
// dmu^2 <= 4, so any call under 3.5e13 pair visits per bin), which leaves 80
// fraction bits; each converted run sum is truncated by less than 2^-80.
constexpr double kFixUnit = 0x1p-80;

// v -> trunc(v 2^80) toward zero as a 128-bit two's-complement integer
// {lo, hi}, from the double's bits without branches: |v| 2^80 = m 2^sh with
// m the 53-bit significand and sh = exponent - 995, so the magnitude is m
// shifted left by sh or right by -sh (the bits shifted out below 2^0 are the
// truncation), then the sign is applied as ~x + 1.  Needs |v| < 2^36
// (sh <= 63); zeros and subnormals have sh = -995 and give 0.
__device__ __forceinline__ void fix_from_double(double v, unsigned long long &lo, unsigned long long &hi) {
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    const int e = static_cast<int>((bits >> 52) & 0x7ffull);
    const unsigned long long m = (bits & 0xfffffffffffffull) | (1ull << 52);
    const int sh = e - 995;
    const int ls = sh > 0 ? sh : 0;                // left shift, 0..63
    const int rs = sh < 0 ? (sh > -63 ? -sh : 63) : 0;  // right shift, 0..63 (m < 2^53: 63 clears it)
    unsigned long long l = (m << ls) >> rs;
    unsigned long long h = (m >> 1) >> (63 - ls);  // m >> (64 - ls), and 0 for ls = 0
    const unsigned long long sgn = static_cast<unsigned long long>(static_cast<long long>(bits) >> 63);
    l ^= sgn;
    h ^= sgn;
    const unsigned long long l1 = l + (sgn & 1ull);
    h += l1 < l ? 1ull : 0ull;
    lo = l1;
    hi = h;
}

// *p += {lo, hi} (p[0] = lo word, p[1] = hi word) with atomics: the carry out
// of the low word is known exactly from its old value, so concurrent adds in
// any order leave the same 128-bit sum.
__device__ __forceinline__ void fix_atomic_add(unsigned long long *p, unsigned long long lo, unsigned long long hi) {
    const unsigned long long old = atomicAdd(&p[0], lo);
    const unsigned long long carry = old + lo < old ? 1ull : 0ull;
    const unsigned long long h = hi + carry;
    if (h) atomicAdd(&p[1], h);
}

__host__ inline double fix_to_double(unsigned long long lo, unsigned long long hi) {
    return static_cast<double>(static_cast<long long>(hi)) * 0x1p-16 + static_cast<double>(lo) * kFixUnit;
}

// Bin storage of one lag: count (u64), sum dmu and sum dmu^2 (128-bit fixed
// point, two u64 words each) = 5 words per bin; a lag's bins are
// [count x nbins][s1 x 2 nbins][s2 x 2 nbins].
constexpr int kBinWords = 5;

// A workgroup's LDS bins hold each fixed-point sum as three words: the sums
// of the addends' bits 0-31 and bits 32-63 (below 2^32 each, so their sums
// are exact while a word takes fewer than 2^32 adds: 256 lanes x fewer than
// 2^24 samples) and of bits 64-127 modulo 2^64.  Adding needs no carry and
// no returned value; lds_fix_total puts the three words back together
// modulo 2^128.  A lag's LDS bins: [count x nbins][s1 x 3 nbins][s2 x 3 nbins].
// A flushed run sum is below 4 x 2^24 = 2^26 in magnitude (a run is shorter
// than the series), inside fix_from_double's range.
constexpr int kLdsBinWords = 7;
constexpr int64_t kMaxDmumuSamples = int64_t{1} << 24;

__device__ __forceinline__ void lds_fix_add(unsigned long long *w, double v) {
    unsigned long long lo, hi;
    fix_from_double(v, lo, hi);
    atomicAdd(&w[0], lo & 0xffffffffull);
    atomicAdd(&w[1], lo >> 32);
    atomicAdd(&w[2], hi);
}

__device__ __forceinline__ void lds_fix_total(const unsigned long long *w, unsigned long long &lo,
                                              unsigned long long &hi) {
    lo = w[0] + (w[1] << 32);
    hi = w[2] + (w[1] >> 32) + (lo < w[0] ? 1ull : 0ull);
}

// Raw buffer loads over one uniform row: the lane's 32-bit byte offset comes
// in a VGPR and a uniform 32-bit offset (a lag's row) in an SGPR, so the
// walk's loads need no per-lane 64-bit address arithmetic.  Descriptor word
// 3 = 0x00020000 (gfx9-family raw access, 32-bit data format); the record
// count admits every offset below 2^31, which the host guarantees.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *row) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return __hiloint2double(static_cast<int>(v[1]), static_cast<int>(v[0]));
}

// kBuf: the finite fast path loads through row_rsrc (every lane and lag
// offset below 2^31 bytes); otherwise through 64-bit addresses.
template <bool kBuf>
__global__ void __launch_bounds__(kTPB)
k_dmumu(const float *__restrict__ mu, int64_t nsave, int64_t n_sel, int64_t ss, int64_t sp,
        const int64_t *__restrict__ pidx, const int64_t *__restrict__ lags, int nlags,
        const float *__restrict__ lo_g, int nbins, float hi_last, int bin_abs, int injection,
        unsigned long long *__restrict__ g_bins, double *__restrict__ part_d, long long *__restrict__ part_i,
        int nbx, int ngy, const int32_t *__restrict__ fin, int lane32) {
    const TileGroup tg = tile_group(ngy);
    if (tg.tile >= nbx) return;  // whole block: before any barrier
    extern __shared__ unsigned char smem[];
    // LDS bins of the group's kGmu lags, laid out as g_bins per lag
    unsigned long long *b_bins = reinterpret_cast<unsigned long long *>(smem);
    float *lo = reinterpret_cast<float *>(b_bins + kGmu * kLdsBinWords * nbins);
    __shared__ Welford sh_w[kTPB / 64];
    __shared__ double sh_d[kTPB / 64];
    __shared__ long long sh_i[kTPB / 64];

    for (int i = threadIdx.x; i < kGmu * kLdsBinWords * nbins; i += kTPB) b_bins[i] = 0ull;
    for (int i = threadIdx.x; i < nbins; i += kTPB) lo[i] = lo_g[i];
    __syncthreads();

    const int gy = tg.group;
    // The last group is padded with its first lag: every load stays in range
    // and the padded slots are never written out (li < nlags below).
    int64_t lag[kGmu];
#pragma unroll
    for (int g = 0; g < kGmu; ++g) {
        const int li = gy * kGmu + g;
        lag[g] = lags[li < nlags ? li : gy * kGmu];
    }
    const int64_t j = static_cast<int64_t>(tg.tile) * kTPB + threadIdx.x;
    const bool active = j < n_sel;

    // Run accumulators of the current start bin (every pair is added; a run
    // whose start was outside the binned range is dropped when it closes),
    // the particle's totals over all valid pairs (sum dmu^2, count), folded
    // in when a run closes, and a stash holding one closed run in range
    // until its LDS flush.  A lane's bin changes rarely, so flushing each
    // run as it closes would run the flush (fixed-point conversions and LDS
    // atomics) for one or two lanes of the wave at a time; with the stash a
    // closed run costs a few register moves, and the wave flushes the
    // stashes of all its lanes together, only when a lane closes a run while
    // its stash is taken.
    unsigned rc[kGmu], pc[kGmu], sc[kGmu];
    double r1[kGmu], r2[kGmu], ps[kGmu], s1[kGmu], s2[kGmu];
#pragma unroll
    for (int g = 0; g < kGmu; ++g) {
        rc[g] = 0;
        pc[g] = 0;
        sc[g] = 0;
        r1[g] = r2[g] = ps[g] = s1[g] = s2[g] = 0.0;
    }
    int cur_bin = -1, st_bin = -1;
    float cur_lo = INFINITY, cur_hi = -INFINITY;
    // first guess of a value's bin for evenly spaced edges
    const float lo0 = lo[0], inv_w = static_cast<float>(nbins) / (hi_last - lo0);
    // starts of the current run visited on the finite fast path, where every
    // lag's pair is valid: their count is added to each rc[g] at the flush
    unsigned run = 0;

    // LDS adds of one run in bin `bin`: pair counts and both fixed-point sums
    auto emit = [&](int bin, const unsigned *c, const double *a1, const double *a2) {
#pragma unroll
        for (int g = 0; g < kGmu; ++g) {
            if (c[g]) {
                unsigned long long *bl = b_bins + g * kLdsBinWords * nbins;
                atomicAdd(&bl[bin], static_cast<unsigned long long>(c[g]));
                lds_fix_add(&bl[nbins + 3 * bin], a1[g]);
                lds_fix_add(&bl[4 * nbins + 3 * bin], a2[g]);
            }
        }
    };
    auto flush_stash = [&]() {
        if (st_bin >= 0) {
            emit(st_bin, sc, s1, s2);
            st_bin = -1;
        }
    };
    // close the current run into the particle's totals and, in range, into
    // the stash (which the caller has emptied)
    auto close_run = [&]() {
        const bool keep = cur_bin >= 0;
#pragma unroll
        for (int g = 0; g < kGmu; ++g) {
            rc[g] += run;
            pc[g] += rc[g];
            ps[g] += r2[g];
            if (keep) {
                sc[g] = rc[g];
                s1[g] = r1[g];
                s2[g] = r2[g];
            }
            rc[g] = 0;
            r1[g] = r2[g] = 0.0;
        }
        if (keep) st_bin = cur_bin;
        run = 0;
    };

    int64_t lag_min = lag[0], lag_max = lag[0], off[kGmu];
#pragma unroll
    for (int g = 0; g < kGmu; ++g) {
        lag_min = lag[g] < lag_min ? lag[g] : lag_min;
        lag_max = lag[g] > lag_max ? lag[g] : lag_max;
        off[g] = lag[g] * ss;
    }
    const int64_t s_end = injection ? (lag_min < nsave ? 1 : 0) : (nsave - lag_min);
    const int64_t s_full = nsave - lag_max;  // below it every lag of the group is in range
    const int64_t s_main = s_end < s_full ? s_end : s_full;
    // Finite fast path: every start below s_fin has a finite mu and finite
    // partners at every lag of the group, so no pair needs a check.  s_fin is
    // the smallest such bound over the wave's particles, so the walk below it
    // advances the wave in step and addresses each sample row as one scalar
    // pointer plus the lane's 32-bit byte offset (lane32: it fits).
    int64_t s_fin = 0;
    if (lane32) {
        int64_t lim = s_main;
        if (active) {
            const int64_t f = static_cast<int64_t>(fin[j]) - lag_max;
            lim = f < lim ? (f > 0 ? f : 0) : lim;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int64_t t = __shfl_xor(lim, o, 64);
            lim = t < lim ? t : lim;
        }
        s_fin = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(lim)));
    }

    if (active) {
        const int64_t col = pidx ? pidx[j] : j;
        const float *row = mu + col * sp;  // sample s of this particle

        // One start sample: bin of mu[s] (a run boundary only when it
        // changes), then the pair with every lag, without branches.  `b[g]`
        // holds mu[s + lag[g]] (any value where `valid[g]` is false).
        auto rebin = [&](float a) {
            const float v = bin_abs ? fabsf(a) : a;
            const bool change = !(v >= cur_lo && v < cur_hi);
            // wave-uniform branches: the common sample, where no lane of the
            // wave changes bin, costs one ballot
            if (!__builtin_amdgcn_ballot_w64(change)) return;
            // a lane must close an in-range run while its stash is taken:
            // every lane with a stash flushes it now
            if (__builtin_amdgcn_ballot_w64(change && cur_bin >= 0 && st_bin >= 0)) flush_stash();
            if (change) {
                close_run();
                if (v < lo[0]) {
                    cur_bin = -1;
                    cur_lo = -INFINITY;
                    cur_hi = lo[0];
                } else if (v >= hi_last) {
                    cur_bin = -1;
                    cur_lo = hi_last;
                    cur_hi = INFINITY;
                } else {
                    // largest b with lo[b] <= v: the guess, else a binary
                    // search on the side of it that holds b
                    int g0 = static_cast<int>((v - lo0) * inv_w);
                    g0 = g0 < 0 ? 0 : (g0 > nbins - 1 ? nbins - 1 : g0);
                    int b0 = 0, b1 = nbins - 1;
                    if (lo[g0] > v)
                        b1 = g0 - 1;
                    else if (g0 + 1 < nbins && lo[g0 + 1] <= v)
                        b0 = g0 + 1;
                    else
                        b0 = b1 = g0;
                    while (b0 < b1) {
                        const int mid = (b0 + b1 + 1) >> 1;
                        if (lo[mid] <= v)
                            b0 = mid;
                        else
                            b1 = mid - 1;
                    }
                    cur_bin = b0;
                    cur_lo = lo[b0];
                    cur_hi = b0 + 1 < nbins ? lo[b0 + 1] : hi_last;
                }
            }
        };
        auto visit = [&](float a, const float *b, const bool *valid) {
            if (!isfinite(a)) return;
            rebin(a);
            const double ad = static_cast<double>(a);
#pragma unroll
            for (int g = 0; g < kGmu; ++g) {
                const bool ok = valid[g] && isfinite(b[g]);
                const double d = ok ? static_cast<double>(b[g]) - ad : 0.0;
                rc[g] += ok ? 1u : 0u;
                r1[g] += d;
                r2[g] = fma(d, d, r2[g]);
            }
        };
        // a and every b[g] finite (start below the particle's first
        // non-finite sample minus the group's largest lag)
        auto visit_finite = [&](float a, const float *b) {
            rebin(a);
            const double ad = static_cast<double>(a);
#pragma unroll
            for (int g = 0; g < kGmu; ++g) {
                const double d = static_cast<double>(b[g]) - ad;
                r1[g] += d;
                r2[g] = fma(d, d, r2[g]);
            }
            ++run;
        };
        bool all[kGmu];
#pragma unroll
        for (int g = 0; g < kGmu; ++g) all[g] = true;
        int64_t s = 0;
        if (s_fin > 0) {
            // three samples in flight in rotating registers
            const uint32_t lane_bytes = static_cast<uint32_t>(col * sp * 4);
            auto at = [&](const float *uniform_row) {
                return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(uniform_row) + lane_bytes);
            };
            uint32_t soff[kGmu];
#pragma unroll
            for (int g = 0; g < kGmu; ++g) soff[g] = static_cast<uint32_t>(off[g] * 4);
            auto load = [&](int64_t t, float &a, float *b) {
                const float *r = mu + t * ss;
                if constexpr (kBuf) {
                    const __amdgpu_buffer_rsrc_t rs = row_rsrc(r);
                    a = buf_load(rs, lane_bytes, 0u, 0.0f);
#pragma unroll
                    for (int g = 0; g < kGmu; ++g) b[g] = buf_load(rs, lane_bytes, soff[g], 0.0f);
                } else {
                    a = at(r);
#pragma unroll
                    for (int g = 0; g < kGmu; ++g) b[g] = at(r + off[g]);
                }
            };
            // Every load is issued unconditionally, past the end clamped to
            // the last start (a row already loaded): the number of loads in
            // flight is then the same on every path, and the compiler waits
            // only for the oldest sample's loads instead of draining them all.
            const int64_t t_last = s_fin - 1;
            auto ahead = [&](int64_t t) { return t < t_last ? t : t_last; };
            float a0, a1, a2, b0[kGmu], b1[kGmu], b2[kGmu];
            load(0, a0, b0);
            load(ahead(1), a1, b1);
            for (;;) {
                load(ahead(s + 2), a2, b2);
                visit_finite(a0, b0);
                if (++s == s_fin) break;
                load(ahead(s + 2), a0, b0);
                visit_finite(a1, b1);
                if (++s == s_fin) break;
                load(ahead(s + 2), a1, b1);
                visit_finite(a2, b2);
                if (++s == s_fin) break;
            }
            row += s * ss;
        }
        // Every lag in range: the loads of sample s+1 are issued before the
        // pairs of sample s are accumulated, so their latency overlaps it.
        if (s < s_main) {
            float an = row[0], bn[kGmu];
#pragma unroll
            for (int g = 0; g < kGmu; ++g) bn[g] = row[off[g]];
            for (; s < s_main; ++s) {
                const float a = an;
                float b[kGmu];
#pragma unroll
                for (int g = 0; g < kGmu; ++g) b[g] = bn[g];
                row += ss;
                if (s + 1 < s_main) {
                    an = row[0];
#pragma unroll
                    for (int g = 0; g < kGmu; ++g) bn[g] = row[off[g]];
                }
                visit(a, b, all);
            }
        }
        for (; s < s_end; ++s, row += ss) {
            float b[kGmu];
            bool valid[kGmu];
#pragma unroll
            for (int g = 0; g < kGmu; ++g) {
                valid[g] = s + lag[g] < nsave;  // uniform
                b[g] = 0.0f;
                if (valid[g]) b[g] = row[off[g]];
            }
            visit(row[0], b, valid);
        }
        flush_stash();
        close_run();
        flush_stash();
    }

#pragma unroll
    for (int g = 0; g < kGmu; ++g) {
        const int li = gy * kGmu + g;
        Welford w{0, 0, 0};
        if (active && pc[g] > 0) {
            w.n = 1.0;
            w.mean = ps[g] / static_cast<double>(pc[g]);
        }
        const Welford bw = block_welford(w, sh_w);
        const double pss = block_sum_d(pc[g] > 0 ? ps[g] : 0.0, sh_d);
        const long long pcs = block_sum_i(static_cast<long long>(pc[g]), sh_i);
        if (threadIdx.x == 0 && li < nlags) {
            const int64_t o = static_cast<int64_t>(li) * nbx + tg.tile;
            part_d[o * 3 + 0] = bw.mean;
            part_d[o * 3 + 1] = bw.m2;
            part_d[o * 3 + 2] = pss;
            part_i[o * 2 + 0] = static_cast<long long>(bw.n);
            part_i[o * 2 + 1] = pcs;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kGmu * nbins; i += kTPB) {
        const int g = i / nbins, b = i % nbins;
        const int li = gy * kGmu + g;
        const unsigned long long *bl = b_bins + g * kLdsBinWords * nbins;
        if (li < nlags && bl[b]) {
            unsigned long long *gl = g_bins + static_cast<int64_t>(li) * kBinWords * nbins;
            atomicAdd(&gl[b], bl[b]);
            unsigned long long l, h;
            lds_fix_total(&bl[nbins + 3 * b], l, h);
            fix_atomic_add(&gl[nbins + 2 * b], l, h);
            lds_fix_total(&bl[4 * nbins + 3 * b], l, h);
            fix_atomic_add(&gl[3 * nbins + 2 * b], l, h);
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long *__restrict__ a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0ull;
}

// fin[j] = index of the first non-finite sample of selected particle j (nsave
// when the whole series is finite): the pair walks need no per-pair checks
// below it.  Threads cover (particle, run of samples) tiles, so the read is
// coalesced for the [save][n] layout and many loads are in flight;
// non-finite samples (rare) lower fin[j] with an atomic min.  fin must hold
// nsave beforehand.
__global__ void __launch_bounds__(256) k_fill_i32(int32_t *__restrict__ a, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

constexpr int kPrefixRun = 16;  // samples per thread of k_finite_prefix

template <class T>
__global__ void __launch_bounds__(256)
k_finite_prefix(const T *__restrict__ mu, int64_t nsave, int64_t n_sel, int64_t ss, int64_t sp,
                const int64_t *__restrict__ pidx, int32_t *__restrict__ fin) {
    // x: particle (coalesced for the [save][n] layout); y: a run of kPrefixRun
    // samples, all loaded before any is tested so the loads overlap
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_sel) return;
    const T *row = mu + (pidx ? pidx[j] : j) * sp;
    for (int64_t s0 = static_cast<int64_t>(blockIdx.y) * kPrefixRun; s0 < nsave;
         s0 += static_cast<int64_t>(gridDim.y) * kPrefixRun) {
        T v[kPrefixRun];
#pragma unroll
        for (int q = 0; q < kPrefixRun; ++q) v[q] = s0 + q < nsave ? row[(s0 + q) * ss] : T(0);
        int first = -1;
#pragma unroll
        for (int q = kPrefixRun - 1; q >= 0; --q)
            if (!isfinite(v[q])) first = q;
        if (first >= 0) {
            atomicMin(&fin[j], static_cast<int32_t>(s0 + first));
            return;  // later runs can only hold larger indices
        }
    }
}

// Fill fin (n_sel values) for the walks; see k_finite_prefix.
template <class T>
void launch_finite_prefix(const T *mu, int64_t nsave, int64_t n_sel, int64_t ss, int64_t sp, const int64_t *pidx,
                          int32_t *fin, hipStream_t st) {
    k_fill_i32<<<static_cast<unsigned>((n_sel + 255) / 256), 256, 0, st>>>(fin, n_sel, static_cast<int32_t>(nsave));
    PT_HIP_CHECK(hipGetLastError());
    const int64_t runs = (nsave + kPrefixRun - 1) / kPrefixRun;
    const dim3 grid(static_cast<unsigned>((n_sel + 255) / 256), static_cast<unsigned>(runs < 65535 ? runs : 65535));
    k_finite_prefix<T><<<grid, 256, 0, st>>>(mu, nsave, n_sel, ss, sp, pidx, fin);
    PT_HIP_CHECK(hipGetLastError());
}

// Merge per-workgroup partials of each lag in workgroup order.
// One workgroup per lag: thread t merges a contiguous run of partials in
// workgroup order, then the fixed tree of block_welford/block_sum_*, so the
// result is deterministic.
__global__ void __launch_bounds__(kTPB)
k_merge_partials(const double *__restrict__ part_d, const long long *__restrict__ part_i, int nlags, int nbx,
                 double *__restrict__ out_d, long long *__restrict__ out_i) {
    __shared__ Welford sh_w[kTPB / 64];
    __shared__ double sh_d[kTPB / 64];
    __shared__ long long sh_i[kTPB / 64];
    const int li = blockIdx.x;
    const int per = (nbx + kTPB - 1) / kTPB;
    const int b0 = threadIdx.x * per, b1 = b0 + per < nbx ? b0 + per : nbx;
    Welford w{0, 0, 0};
    double pss = 0.0;
    long long pcs = 0;
    for (int b = b0; b < b1; ++b) {
        const int64_t o = static_cast<int64_t>(li) * nbx + b;
        Welford x{static_cast<double>(part_i[o * 2]), part_d[o * 3], part_d[o * 3 + 1]};
        w = chan(w, x);
        pss += part_d[o * 3 + 2];
        pcs += part_i[o * 2 + 1];
    }
    const Welford bw = block_welford(w, sh_w);
    const double ps = block_sum_d(pss, sh_d);
    const long long pc = block_sum_i(pcs, sh_i);
    if (threadIdx.x == 0 && li < nlags) {
        out_d[li * 3 + 0] = bw.mean;
        out_d[li * 3 + 1] = bw.m2;
        out_d[li * 3 + 2] = ps;
        out_i[li * 2 + 0] = static_cast<long long>(bw.n);
        out_i[li * 2 + 1] = pc;
    }
}

// ------------------------------------------------------------------ D_pp ---

__device__ __forceinline__ double pmag(const float *__restrict__ mom, int64_t s, int64_t total,
                                       int64_t col) {
    const float *q = mom + s * 3 * total + col;
    const double x = q[0], y = q[total], z = q[2 * total];
    return sqrt(x * x + y * y + z * z);
}
// |p| read from a precomputed FP64 cache [save][total].
__device__ __forceinline__ double pmag(const double *__restrict__ pm, int64_t s, int64_t total, int64_t col) {
    return pm[s * total + col];
}

// |p| in FP64 of every (sample, selected particle): momentum_magnitude
// (compute_dpp_full.jl:232-234) evaluated once per sample, not once per pair.
__global__ void k_pmag(const float *__restrict__ mom, int64_t nsave, int64_t total, int64_t c0, int64_t nc,
                       const int64_t *__restrict__ pidx, double *__restrict__ out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nsave * nc) return;
    const int64_t s = t / nc, jj = t % nc;
    const int64_t col = pidx ? pidx[c0 + jj] : c0 + jj;
    out[t] = pmag(mom, s, total, col);
}

// D_pp pairs over a series [save][total]: the walk of k_dmumu without bins.
// Per lag it sums d = series(s+lag) - series(s) and d^2 over the valid pairs.
// T = double, kDelta = false: the FP64 |p| series, d = Delta p; the host forms
// the normalised sums as (sum d)/p0 and (sum d^2)/p0^2, the reference's
// sum(Delta p/p0) and sum((Delta p/p0)^2) up to the rounding of the terms
// (two FP64 accumulators per lag fewer in the walk).
// T = float, kDelta = true: a relative series delta = |p|/p_ref - 1 (d exact
// in FP64); the host scales the sums by p_ref and p_ref/p0.
template <class T, bool kDelta, bool kBuf>
__global__ void __launch_bounds__(kTPB)
k_dpp(const T *__restrict__ pm, int64_t nsave, int64_t total, int64_t n_sel,
      const int64_t *__restrict__ pidx, const int64_t *__restrict__ lags, int nlags,
      double *__restrict__ part_d, long long *__restrict__ part_i, int nbx, int ngy,
      const int32_t *__restrict__ fin, int lane32) {
    constexpr int kG = kDppLags<T>;
    const TileGroup tg = tile_group(ngy);
    if (tg.tile >= nbx) return;
    __shared__ double sh_d[kTPB / 64];
    __shared__ long long sh_i[kTPB / 64];
    const int gy = tg.group;
    int64_t lag[kG];
#pragma unroll
    for (int g = 0; g < kG; ++g) {
        const int li = gy * kG + g;
        lag[g] = lags[li < nlags ? li : gy * kG];
    }
    const int64_t j = static_cast<int64_t>(tg.tile) * kTPB + threadIdx.x;
    unsigned cnt[kG];
    double s1[kG], s2[kG];
#pragma unroll
    for (int g = 0; g < kG; ++g) {
        cnt[g] = 0;
        s1[g] = s2[g] = 0.0;
    }
    int64_t lag_min = lag[0], lag_max = lag[0], off[kG];
#pragma unroll
    for (int g = 0; g < kG; ++g) {
        lag_min = lag[g] < lag_min ? lag[g] : lag_min;
        lag_max = lag[g] > lag_max ? lag[g] : lag_max;
        off[g] = lag[g] * total;
    }
    const int64_t s_end = nsave - lag_min, s_full = nsave - lag_max;
    const int64_t s_main = s_end < s_full ? s_end : s_full;
    // Finite fast path below the wave's smallest s_fin, as in k_dmumu.
    int64_t s_fin = 0;
    if (lane32) {
        int64_t lim = s_main;
        if (j < n_sel) {
            const int64_t f = static_cast<int64_t>(fin[j]) - lag_max;
            lim = f < lim ? (f > 0 ? f : 0) : lim;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int64_t t = __shfl_xor(lim, o, 64);
            lim = t < lim ? t : lim;
        }
        s_fin = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(lim)));
    }
    if (j < n_sel) {
        const int64_t col = pidx ? pidx[j] : j;
        const T *row = pm + col;
        // `pe[g]` = the series at s + lag[g] (any value where `valid[g]` is false)
        auto visit = [&](double pst, const double *pe, const bool *valid) {
            if (!isfinite(pst)) return;
#pragma unroll
            for (int g = 0; g < kG; ++g) {
                const bool ok = valid[g] && isfinite(pe[g]);
                const double d = ok ? pe[g] - pst : 0.0;
                cnt[g] += ok ? 1u : 0u;
                s1[g] += d;
                s2[g] = fma(d, d, s2[g]);
            }
        };
        bool all[kG];
#pragma unroll
        for (int g = 0; g < kG; ++g) all[g] = true;
        int64_t s = 0;
        if (s_fin > 0) {
            // every pair valid: the count is the number of starts
            const uint32_t lane_bytes = static_cast<uint32_t>(col * static_cast<int64_t>(sizeof(T)));
            // samples stay in their stored type until used (a float series
            // holds half the registers of its FP64 values)
            auto at = [&](const T *uniform_row) {
                return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(uniform_row) + lane_bytes);
            };
            uint32_t soff[kG];
#pragma unroll
            for (int g = 0; g < kG; ++g) soff[g] = static_cast<uint32_t>(off[g] * static_cast<int64_t>(sizeof(T)));
            auto load = [&](int64_t t, T &a, T *b) {
                const T *r = pm + t * total;
                if constexpr (kBuf) {
                    const __amdgpu_buffer_rsrc_t rs = row_rsrc(r);
                    a = buf_load(rs, lane_bytes, 0u, T{});
#pragma unroll
                    for (int g = 0; g < kG; ++g) b[g] = buf_load(rs, lane_bytes, soff[g], T{});
                } else {
                    a = at(r);
#pragma unroll
                    for (int g = 0; g < kG; ++g) b[g] = at(r + off[g]);
                }
            };
            auto visit_finite = [&](T st, const T *pe) {
                const double pst = static_cast<double>(st);
#pragma unroll
                for (int g = 0; g < kG; ++g) {
                    const double d = static_cast<double>(pe[g]) - pst;
                    s1[g] += d;
                    s2[g] = fma(d, d, s2[g]);
                }
            };
            // three samples in flight in rotating registers, every load
            // issued (clamped to the last start) as in k_dmumu
            const int64_t t_last = s_fin - 1;
            auto ahead = [&](int64_t t) { return t < t_last ? t : t_last; };
            T a0, a1, a2, b0[kG], b1[kG], b2[kG];
            load(0, a0, b0);
            load(ahead(1), a1, b1);
            for (;;) {
                load(ahead(s + 2), a2, b2);
                visit_finite(a0, b0);
                if (++s == s_fin) break;
                load(ahead(s + 2), a0, b0);
                visit_finite(a1, b1);
                if (++s == s_fin) break;
                load(ahead(s + 2), a1, b1);
                visit_finite(a2, b2);
                if (++s == s_fin) break;
            }
#pragma unroll
            for (int g = 0; g < kG; ++g) cnt[g] += static_cast<unsigned>(s);
            row += s * total;
        }
        // every lag in range: sample s+1 is loaded before sample s is used
        if (s < s_main) {
            T an = row[0], bn[kG];
#pragma unroll
            for (int g = 0; g < kG; ++g) bn[g] = row[off[g]];
            for (; s < s_main; ++s) {
                const double a = an;
                double b[kG];
#pragma unroll
                for (int g = 0; g < kG; ++g) b[g] = bn[g];
                row += total;
                if (s + 1 < s_main) {
                    an = row[0];
#pragma unroll
                    for (int g = 0; g < kG; ++g) bn[g] = row[off[g]];
                }
                visit(a, b, all);
            }
        }
        for (; s < s_end; ++s, row += total) {
            double b[kG];
            bool valid[kG];
#pragma unroll
            for (int g = 0; g < kG; ++g) {
                valid[g] = s + lag[g] < nsave;  // uniform
                b[g] = 0.0;
                if (valid[g]) b[g] = row[off[g]];
            }
            visit(row[0], b, valid);
        }
    }
#pragma unroll
    for (int g = 0; g < kG; ++g) {
        const int li = gy * kG + g;
        const long long c = block_sum_i(static_cast<long long>(cnt[g]), sh_i);
        const double a = block_sum_d(s1[g], sh_d);
        const double b = block_sum_d(s2[g], sh_d);
        if (threadIdx.x == 0 && li < nlags) {
            const int64_t o = static_cast<int64_t>(li) * nbx + tg.tile;
            part_i[o] = c;
            part_d[o * 2 + 0] = a;
            part_d[o * 2 + 1] = b;
        }
    }
}

// D_pp walk over LDS tiles.  A workgroup stages the whole series of kP
// selected particles in LDS (rows [save][kP] of T, read from HBM once) and
// walks every lag over them: wave w takes lag groups w, w + 4, ...; lane
// (k, p) = (lane / kP, lane % kP) walks the k-th of 64 / kP runs of
// consecutive starts of particle p.  The run length is odd, so the lanes of a
// read fall on distinct LDS banks.  (k_dpp above re-reads each sample from
// the L2 once per lag; here the L2 sees it once per workgroup.)  A workgroup
// takes tiles first, first + gridDim.x, ... (first: below) and adds each
// lag's wave sum into its own slot part_*[lag * gridDim.x + blockIdx.x] in
// tile order, so the result does not depend on scheduling.  Tiles whose particles are finite at
// every sample skip the finiteness tests; their counts are np * (nsave - lag).
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Lags per wave pass of k_dpp_tile: with eight, the float series' walk held
// 134 VGPRs (three waves per SIMD) and ran 15% slower; the FP64 walk is bound by
// LDS bandwidth either way (five reads of eight bytes per four pairs).
constexpr int kTileLags = 4;

// kMom: the source is FP32 momenta [save][3][total] and the tile holds their
// FP64 |p| (momentum_magnitude, evaluated while staging); otherwise the
// source is the series itself, [save][total] of T.
template <class T, bool kDelta, int kP, bool kMom>
__global__ void __launch_bounds__(kTPB)
k_dpp_tile(const void *__restrict__ src, int32_t nsave, int64_t total, int64_t n_sel,
           const int64_t *__restrict__ pidx, const int64_t *__restrict__ lags, int nlags,
           double *__restrict__ part_d, long long *__restrict__ part_i) {
    static_assert(!kMom || (std::is_same<T, double>::value && !kDelta), "momenta stage an FP64 |p| series");
    constexpr int kG = kTileLags;
    constexpr int kK = 64 / kP;
    static_assert(kP >= 1 && kP <= 64 && (kP & (kP - 1)) == 0, "kP: a power of two up to 64");
    extern __shared__ __align__(16) unsigned char smem[];
    T *rows = reinterpret_cast<T *>(smem);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = lane & (kP - 1), k = lane / kP;
    const int ngroups = (nlags + kG - 1) / kG;
    const int64_t ntiles = (n_sel + kP - 1) / kP;
    const int nel = nsave * kP;
    // Consecutive tiles share cache lines (a row of kP < 16 FP64 values is
    // part of a line), so they run on one XCD at the same time: within each
    // run of 32 workgroups, the four on XCD x (= blockIdx % 8) take tiles
    // 4x .. 4x + 3.  A bijection on [0, gridDim.x).
    const int b = static_cast<int>(blockIdx.x);
    const int full = static_cast<int>(gridDim.x) & ~31;
    const int first = b < full ? (b & ~31) + 4 * (b & 7) + ((b & 31) >> 3) : b;
    for (int64_t tile = first; tile < ntiles; tile += gridDim.x) {
        const int64_t j0 = tile * kP;
        const int np = n_sel - j0 < kP ? static_cast<int>(n_sel - j0) : kP;
        __syncthreads();  // the previous tile's walk is done with the rows
        // stage: kU samples in flight per thread before their LDS writes
        // (momenta: three loads each); the tile is finite when every staged
        // sample is
        constexpr int kU = kMom ? 4 : 8;
        bool ok = true;
        for (int e0 = threadIdx.x; e0 < nel; e0 += kU * kTPB) {
            T v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int e = e0 + u * kTPB;
                const int q = e & (kP - 1);
                v[u] = T(0);
                if (e < nel && q < np) {
                    const int64_t col = pidx ? pidx[j0 + q] : j0 + q;
                    if constexpr (kMom)
                        v[u] = pmag(static_cast<const float *>(src), e / kP, total, col);
                    else
                        v[u] = static_cast<const T *>(src)[static_cast<int64_t>(e / kP) * total + col];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (e0 + u * kTPB < nel) {
                    rows[e0 + u * kTPB] = v[u];
                    ok = ok && isfinite(v[u]);
                }
        }
        const bool finite = __syncthreads_and(ok) != 0;
        const T *col = rows + p;
        for (int gy = wave; gy < ngroups; gy += kTPB / 64) {
            int32_t off[kG];
            bool valid[kG];
            int32_t lmin = nsave, lmax = 0;
#pragma unroll
            for (int g = 0; g < kG; ++g) {
                const int li = gy * kG + g;
                const int64_t lv = li < nlags ? lags[li] : nsave;
                valid[g] = lv < nsave;
                off[g] = valid[g] ? static_cast<int32_t>(lv) : 0;
                if (valid[g]) {
                    lmin = off[g] < lmin ? off[g] : lmin;
                    lmax = off[g] > lmax ? off[g] : lmax;
                }
            }
            double s1[kG], s2[kG];
            unsigned cnt[kG];
#pragma unroll
            for (int g = 0; g < kG; ++g) {
                s1[g] = s2[g] = 0.0;
                cnt[g] = 0;
            }
            if (lmin < nsave) {
                // starts [0, nsave - lmin); below nsave - lmax every lag is in range
                const int s_end = nsave - lmin, s_main = nsave - lmax;
                const int seg = ((s_end + kK - 1) / kK) | 1;
                const int s_lo = k * seg < s_end ? k * seg : s_end;
                const int s_hi = s_lo + seg < s_end ? s_lo + seg : s_end;
                const int m_hi = s_hi < s_main ? s_hi : s_main;
                const int t_lo = s_lo > s_main ? s_lo : s_main;
                if (finite) {
                    // a lag out of range pairs the start with itself: d = 0
#pragma unroll 2
                    for (int s = s_lo; s < m_hi; ++s) {
                        const double a = static_cast<double>(col[s * kP]);
#pragma unroll
                        for (int g = 0; g < kG; ++g) {
                            const double d = static_cast<double>(col[(s + off[g]) * kP]) - a;
                            s1[g] += d;
                            s2[g] = fma(d, d, s2[g]);
                        }
                    }
                    for (int s = t_lo; s < s_hi; ++s) {
                        const double a = static_cast<double>(col[s * kP]);
#pragma unroll
                        for (int g = 0; g < kG; ++g) {
                            const int e = s + off[g] < nsave ? s + off[g] : s;
                            const double d = static_cast<double>(col[e * kP]) - a;
                            s1[g] += d;
                            s2[g] = fma(d, d, s2[g]);
                        }
                    }
                } else {
                    const bool live = p < np;
                    for (int s = s_lo; s < s_hi; ++s) {
                        const T av = col[s * kP];
                        const bool fa = live && isfinite(av);
#pragma unroll
                        for (int g = 0; g < kG; ++g) {
                            const bool in = valid[g] && s + off[g] < nsave;
                            const T bv = col[(in ? s + off[g] : s) * kP];
                            const bool ok = fa && in && isfinite(bv);
                            const double d = ok ? static_cast<double>(bv) - static_cast<double>(av) : 0.0;
                            cnt[g] += ok ? 1u : 0u;
                            s1[g] += d;
                            s2[g] = fma(d, d, s2[g]);
                        }
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < kG; ++g) {
                const int li = gy * kG + g;
                const double a = wave_sum_d(s1[g]);
                const double b = wave_sum_d(s2[g]);
                long long c;
                if (finite) {
                    c = valid[g] ? static_cast<long long>(np) * (nsave - off[g]) : 0;
                } else {
                    c = static_cast<long long>(cnt[g]);
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
                }
                if (lane == 0 && li < nlags) {
                    const int64_t o = static_cast<int64_t>(li) * gridDim.x + blockIdx.x;
                    if (tile == first) {
                        part_d[o * 2 + 0] = a;
                        part_d[o * 2 + 1] = b;
                        part_i[o] = c;
                    } else {
                        part_d[o * 2 + 0] += a;
                        part_d[o * 2 + 1] += b;
                        part_i[o] += c;
                    }
                }
            }
        }
    }
}

// Particles per LDS tile of k_dpp_tile for a series of nsave samples of
// `elem` bytes: the largest power of two up to 16 whose tile stays within
// 40 KiB (four workgroups per CU), else up to 60 KiB (with the kernel's
// static LDS, inside a workgroup's default 64 KiB); 0 when one particle's
// series does not fit (the global walk k_dpp runs instead).
inline int dpp_tile_particles(int64_t nsave, size_t elem) {
    const int64_t row = nsave * static_cast<int64_t>(elem);
    for (int64_t cap : {int64_t{40} << 10, int64_t{60} << 10})
        for (int kp = 16; kp >= 1; kp >>= 1)
            if (row * kp <= cap) return kp;
    return 0;
}

constexpr int kDppTileBlocks = 2048;  // workgroups of a k_dpp_tile launch at most

__global__ void __launch_bounds__(kTPB)
k_merge_dpp(const double *__restrict__ part_d, const long long *__restrict__ part_i, int nlags, int nbx,
            double *__restrict__ out_d, long long *__restrict__ out_i) {
    __shared__ double sh_d[kTPB / 64];
    __shared__ long long sh_i[kTPB / 64];
    const int li = blockIdx.x;
    const int per = (nbx + kTPB - 1) / kTPB;
    const int b0 = threadIdx.x * per, b1 = b0 + per < nbx ? b0 + per : nbx;
    double a = 0, b = 0;
    long long n = 0;
    for (int k = b0; k < b1; ++k) {
        const int64_t o = static_cast<int64_t>(li) * nbx + k;
        n += part_i[o];
        a += part_d[o * 2 + 0];
        b += part_d[o * 2 + 1];
    }
    a = block_sum_d(a, sh_d);
    b = block_sum_d(b, sh_d);
    n = block_sum_i(n, sh_i);
    if (threadIdx.x == 0 && li < nlags) {
        out_i[li] = n;
        out_d[li * 2 + 0] = a;
        out_d[li * 2 + 1] = b;
    }
}

// kinetic_energy_gev_from_pmag (compute_dpp_full.jl:236-240) at snapshot saves.
__global__ void k_energy(const float *__restrict__ mom, int64_t total, int64_t n_sel,
                         const int64_t *__restrict__ pidx, const int64_t *__restrict__ snap, int64_t k,
                         double *__restrict__ out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_sel * k) return;
    const int64_t jj = t / k, q = t % k;
    const int64_t col = pidx ? pidx[jj] : jj;
    const double pm = pmag(mom, snap[q], total, col);
    double e = NAN;
    if (isfinite(pm)) {
        const double r = pm / (kMp * kC);
        const double gamma = sqrt(1.0 + r * r);
        e = (gamma - 1.0) * kMp * kC * kC / kGeVtoJ;
    }
    out[q + k * jj] = e;
}

// ----------------------------------------------------- mu from phase space -

// trilinear_sample (phase_space_gpu_runner.jl:247-276) in compute precision
// T, split in two: the 1-based lower corner (i0, j0, k0), clamped as the
// reference clamps it, with the fractions (tx, ty, tz); and the
// interpolation of the cell's 8 corner values (corner ci = dx + 2 dy + 4 dz).
template <class T>
struct PlanarGeom {
    T lo[3], hi[3], d[3];
    int64_t n[3];
};

template <class T>
__device__ __forceinline__ void trilinear_cell(T x, T y, T z, const PlanarGeom<T> &g, int64_t ijk[3], T t[3]) {
    const T f[3] = {(x - g.lo[0]) / g.d[0] + T(1), (y - g.lo[1]) / g.d[1] + T(1), (z - g.lo[2]) / g.d[2] + T(1)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T fl = floor(f[c]);
        const int64_t hi = g.n[c] - 1;
        ijk[c] = fl < T(1) ? 1 : (fl > static_cast<T>(hi) ? hi : static_cast<int64_t>(fl));
        t[c] = f[c] - static_cast<T>(ijk[c]);
    }
}

template <class T>
__device__ __forceinline__ T trilinear_corners(const T v[8], const T t[3]) {
    const T tx = t[0], ty = t[1], tz = t[2];
    const T c00 = v[0] * (1 - tx) + v[1] * tx;
    const T c10 = v[2] * (1 - tx) + v[3] * tx;
    const T c01 = v[4] * (1 - tx) + v[5] * tx;
    const T c11 = v[6] * (1 - tx) + v[7] * tx;
    const T c0 = c00 * (1 - ty) + c10 * ty;
    const T c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

// mu for the samples [s0, s0 + seg) of one particle (reconstruct_mu_kernel!,
// compute_delta_mu2_curve.jl:151-184).  A thread walks one particle's
// samples, so consecutive samples mostly fall in the same cell: the 24 corner
// values of Bx, By, Bz stay in registers and are reloaded only when the
// cell changes.  Lanes of a wave hold consecutive particles, so the position
// and momentum reads ([s][3][chunk]) are coalesced, as are the mu writes in
// the sample-major layout ([s][chunk]; layout 1); layout 0 is the
// reference's mu_chunk, Julia (nsave, chunk), each particle's series
// contiguous.
template <class T, class S, class O>
__global__ void __launch_bounds__(256) k_mu_from_phase(const S *__restrict__ pos, const S *__restrict__ mom,
                                                       int64_t chunk, int64_t nsave, int32_t seg,
                                                       const T *__restrict__ F, PlanarGeom<T> g,
                                                       O *__restrict__ mu, int sample_major) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nseg = (nsave + seg - 1) / seg;
    if (t >= chunk * nseg) return;
    const int64_t p = t % chunk;
    const int64_t s0 = (t / chunk) * seg;
    const int64_t s1 = s0 + seg < nsave ? s0 + seg : nsave;
    const int64_t nn = g.n[0] * g.n[1] * g.n[2];
    const int64_t sy = g.n[0], sz = g.n[0] * g.n[1];
    int64_t cur[3] = {-1, -1, -1};
    T v[3][8];
    for (int64_t s = s0; s < s1; ++s) {
        const T x = static_cast<T>(pos[(s * 3 + 0) * chunk + p]);
        const T y = static_cast<T>(pos[(s * 3 + 1) * chunk + p]);
        const T z = static_cast<T>(pos[(s * 3 + 2) * chunk + p]);
        const T px = static_cast<T>(mom[(s * 3 + 0) * chunk + p]);
        const T py = static_cast<T>(mom[(s * 3 + 1) * chunk + p]);
        const T pz = static_cast<T>(mom[(s * 3 + 2) * chunk + p]);
        T value = T(NAN);
        if (!(isnan(x) || isnan(y) || isnan(z) || isnan(px) || isnan(py) || isnan(pz)) && g.lo[0] <= x &&
            x <= g.hi[0] && g.lo[1] <= y && y <= g.hi[1] && g.lo[2] <= z && z <= g.hi[2]) {
            int64_t ijk[3];
            T tt[3];
            trilinear_cell(x, y, z, g, ijk, tt);
            if (ijk[0] != cur[0] || ijk[1] != cur[1] || ijk[2] != cur[2]) {
                const int64_t base = (ijk[0] - 1) + g.n[0] * ((ijk[1] - 1) + g.n[1] * (ijk[2] - 1));
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T *a = F + c * nn + base;
#pragma unroll
                    for (int ci = 0; ci < 8; ++ci)
                        v[c][ci] = a[((ci & 1) ? 1 : 0) + ((ci & 2) ? sy : 0) + ((ci & 4) ? sz : 0)];
                }
                cur[0] = ijk[0];
                cur[1] = ijk[1];
                cur[2] = ijk[2];
            }
            const T bx = trilinear_corners(v[0], tt);
            const T by = trilinear_corners(v[1], tt);
            const T bz = trilinear_corners(v[2], tt);
            // pitch_mu_from_pB, compute_delta_mu2_curve.jl:101-110
            const T pn = sqrt(px * px + py * py + pz * pz);
            const T bn = sqrt(bx * bx + by * by + bz * bz);
            if (isfinite(pn) && isfinite(bn) && pn != T(0) && bn != T(0)) {
                T m = (px * bx + py * by + pz * bz) / (pn * bn);
                value = m < T(-1) ? T(-1) : (m > T(1) ? T(1) : m);
            }
        }
        mu[sample_major ? s * chunk + p : s + nsave * p] = static_cast<O>(value);
    }
}

template <class F>
int guarded(F &&fn) {
    try {
        fn();
        return PT_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (...) {
        set_error("unknown C++ exception");
        return PT_ERR_STATE;
    }
}

// A device buffer for the duration of one call, from the context's pool.
struct DevBuf {
    pt_ctx *ctx;
    void *p = nullptr;
    size_t granted = 0;
    explicit DevBuf(pt_ctx *c) : ctx(c) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { pool_put(ctx, p, granted); }
    void *alloc(size_t bytes) {
        pool_put(ctx, p, granted);
        p = nullptr;
        p = pool_get(ctx, bytes ? bytes : 1, &granted);
        return p;
    }
};

// Bring a host or device array into device memory (owned copy in `hold`).

template <class T>
const T *to_device(const T *src, int on_device, size_t count, DevBuf &hold, hipStream_t s) {
    if (on_device || count == 0) return src;
    void *d = hold.alloc(count * sizeof(T));
    PT_HIP_CHECK(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return static_cast<const T *>(d);
}

float ceil_to_float(double e) {
    float f = static_cast<float>(e);
    if (static_cast<double>(f) < e) f = std::nextafter(f, INFINITY);
    return f;
}
float floor_to_float(double e) {
    float f = static_cast<float>(e);
    if (static_cast<double>(f) > e) f = std::nextafter(f, -INFINITY);
    return f;
}

void combine_welford_host(int64_t *counts, double *means, double *m2s, int64_t li, int64_t lc,
                          double lmean, double lm2) {
    // combine_welford_stats!, compute_delta_mu2_dmumu_full.jl:439-455
    if (lc <= 0) return;
    const int64_t gc = counts[li];
    if (gc == 0) {
        counts[li] = lc;
        means[li] = lmean;
        m2s[li] = lm2;
        return;
    }
    const int64_t cc = gc + lc;
    const double delta = lmean - means[li];
    means[li] += delta * static_cast<double>(lc) / static_cast<double>(cc);
    m2s[li] += lm2 + delta * delta * static_cast<double>(gc) * static_cast<double>(lc) / static_cast<double>(cc);
    counts[li] = cc;
}

// ------------------------------------------------ D_mumu call plumbing ---
//
// One call: validate and derive the bin thresholds (dmumu_plan), place the
// device work areas (dmumu_carve), run the kernels (dmumu_enqueue), copy the
// results back (dmumu_download) and merge them into the caller's
// accumulators (dmumu_merge).  pt_dmumu_accumulate takes its device memory
// from the context's pool; pt_dmumu_accumulate_noalloc from the library's
// fixed scratch array.

constexpr int64_t kMaxLagsPerCall = 1 << 20;

struct DmumuPlan {
    int64_t nsave, n_sel, ss, sp, nlags, nbins;
    int64_t max_col;
    size_t mu_count;         // extent of mu the kernels may touch (elements)
    int nbx, ngy, lane32, buf;
    const int64_t *lags;     // caller's host lag steps
    float lo[kMaxBins];      // float thresholds equivalent to the FP64 edges
    float hi_last;
    size_t bins_words, parts;
    size_t work_bytes;       // bins + partials + merged partials
};

struct DmumuDev {
    const float *mu;
    const int64_t *idx, *lags;
    const float *lo;
    unsigned long long *bins;
    double *part_d, *out_d;
    long long *part_i, *out_i;
    int32_t *fin;
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

// Selected series gathered sample-major: dst[s * n_sel + j] = the sample s of
// selected particle j (src column idx[j], or j), from a source whose series
// are contiguous (stride_s = 1, stride_p between particles).  A block moves
// a tile of 32 particles x 64 samples through LDS, so the reads (64
// consecutive samples of one particle per wave) and the writes (32
// consecutive particles per half-wave) are both coalesced.
__global__ void __launch_bounds__(256) k_series_sample_major(const float *__restrict__ src, int64_t nsave,
                                                             int64_t n_sel, int64_t sp,
                                                             const int64_t *__restrict__ idx,
                                                             float *__restrict__ dst) {
    __shared__ float tile[32][65];
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * 32, s0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int t = static_cast<int>(threadIdx.x);
    const int ls = t & 63, lp = t >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int pj = lp + 4 * k;
        const int64_t j = j0 + pj, s = s0 + ls;
        float v = 0.0f;
        if (j < n_sel && s < nsave) v = src[(idx ? idx[j] : j) * sp + s];
        tile[pj][ls] = v;
    }
    __syncthreads();
    const int wp = t & 31, ws = t >> 5;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int si = ws + 8 * k;
        const int64_t j = j0 + wp, s = s0 + si;
        if (j < n_sel && s < nsave) dst[s * n_sel + j] = tile[wp][si];
    }
}

DmumuPlan dmumu_plan(pt_ctx *ctx, const float *mu, int64_t nsave, int64_t n_sel, int64_t stride_s,
                     int64_t stride_p, const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags,
                     const double *mu_edges, int64_t n_bins, int start_mode) {
    PT_REQUIRE(ctx && mu && lag_steps && mu_edges, "pt_dmumu_accumulate: null argument");
    PT_REQUIRE(nsave >= 1 && n_sel >= 0 && nlags >= 1, "pt_dmumu_accumulate: empty input");
    PT_REQUIRE(nsave < kMaxDmumuSamples, "pt_dmumu_accumulate: series longer than 2^24 samples");
    PT_REQUIRE(n_bins >= 1 && n_bins <= kMaxBins, "pt_dmumu_accumulate: n_bins must be in [1, 256]");
    PT_REQUIRE(start_mode == 0 || start_mode == 1, "Unknown dmumu_start_mode");
    PT_REQUIRE(nlags <= kMaxLagsPerCall, "pt_dmumu_accumulate: too many lags");
    for (int64_t q = 0; q < nlags; ++q) PT_REQUIRE(lag_steps[q] >= 1, "pt_dmumu_accumulate: lag steps must be >= 1");
    for (int64_t b = 0; b < n_bins; ++b)
        PT_REQUIRE(mu_edges[b] < mu_edges[b + 1], "pt_dmumu_accumulate: edges must increase");
    DmumuPlan P;
    P.nsave = nsave;
    P.n_sel = n_sel;
    P.ss = stride_s;
    P.sp = stride_p;
    P.nlags = nlags;
    P.nbins = n_bins;
    P.lags = lag_steps;
    for (int64_t b = 0; b < n_bins; ++b) P.lo[b] = ceil_to_float(mu_edges[b]);
    P.hi_last = std::nextafter(floor_to_float(mu_edges[n_bins]), INFINITY);
    // extent of the mu array the kernels may touch, for host staging
    P.max_col = n_sel > 0 ? n_sel - 1 : 0;
    if (particle_index) {
        P.max_col = 0;
        for (int64_t j = 0; j < n_sel; ++j) {
            PT_REQUIRE(particle_index[j] >= 0, "pt_dmumu_accumulate: negative particle index");
            P.max_col = std::max(P.max_col, particle_index[j]);
        }
    }
    P.mu_count = static_cast<size_t>((nsave - 1) * stride_s + P.max_col * stride_p + 1);
    P.nbx = static_cast<int>((n_sel + kTPB - 1) / kTPB);
    P.ngy = static_cast<int>((nlags + kGmu - 1) / kGmu);
    // the fast path's 32-bit lane offsets (bytes) cover every particle
    P.lane32 = P.max_col * stride_p < (int64_t{1} << 30) && nsave < (int64_t{1} << 31) ? 1 : 0;
    // buffer loads: the lane's and the largest lag's byte offsets below 2^31
    int64_t lag_top = 0;
    for (int64_t q = 0; q < nlags; ++q) lag_top = std::max(lag_top, std::min(lag_steps[q], nsave));
    P.buf = P.lane32 && (P.max_col * stride_p + lag_top * stride_s) < (int64_t{1} << 29) ? 1 : 0;
    const size_t nl = static_cast<size_t>(nlags), nb = static_cast<size_t>(n_bins);
    P.bins_words = nl * nb * kBinWords;
    P.parts = nl * static_cast<size_t>(P.nbx);
    P.work_bytes = align256(P.bins_words * 8) + align256(P.parts * 24) + align256(P.parts * 16) +
                   align256(nl * 24) + align256(nl * 16);
    return P;
}

void dmumu_carve(const DmumuPlan &P, char *base, DmumuDev &D) {
    const size_t nl = static_cast<size_t>(P.nlags);
    size_t off = 0;
    D.bins = reinterpret_cast<unsigned long long *>(base + off);
    off += align256(P.bins_words * 8);
    D.part_d = reinterpret_cast<double *>(base + off);
    off += align256(P.parts * 24);
    D.part_i = reinterpret_cast<long long *>(base + off);
    off += align256(P.parts * 16);
    D.out_d = reinterpret_cast<double *>(base + off);
    off += align256(nl * 24);
    D.out_i = reinterpret_cast<long long *>(base + off);
}

void dmumu_enqueue(const DmumuPlan &P, const DmumuDev &D, int bin_abs, int start_mode, hipStream_t st) {
    // the bins start at zero; a fill kernel rather than a memset node keeps
    // the call valid inside a captured graph
    const int64_t nw = static_cast<int64_t>(P.bins_words);
    k_fill_u64<<<static_cast<unsigned>((nw + 255) / 256), 256, 0, st>>>(D.bins, nw);
    PT_HIP_CHECK(hipGetLastError());
    launch_finite_prefix(D.mu, P.nsave, P.n_sel, P.ss, P.sp, D.idx, D.fin, st);
    const size_t shm = static_cast<size_t>(kGmu) * static_cast<size_t>(P.nbins) * kLdsBinWords * 8 +
                       static_cast<size_t>(P.nbins) * 4;
    auto *kern = P.buf ? k_dmumu<true> : k_dmumu<false>;
    kern<<<tile_group_blocks(P.nbx, P.ngy), kTPB, shm, st>>>(
        D.mu, P.nsave, P.n_sel, P.ss, P.sp, D.idx, D.lags, static_cast<int>(P.nlags), D.lo,
        static_cast<int>(P.nbins), P.hi_last, bin_abs ? 1 : 0, start_mode, D.bins, D.part_d, D.part_i, P.nbx,
        P.ngy, D.fin, P.lane32);
    PT_HIP_CHECK(hipGetLastError());
    k_merge_partials<<<static_cast<unsigned>(P.nlags), kTPB, 0, st>>>(D.part_d, D.part_i,
                                                                      static_cast<int>(P.nlags), P.nbx, D.out_d,
                                                                      D.out_i);
    PT_HIP_CHECK(hipGetLastError());
}

void dmumu_download(const DmumuPlan &P, const DmumuDev &D, unsigned long long *hb, double *hd, long long *hi,
                    hipStream_t st) {
    const size_t nl = static_cast<size_t>(P.nlags);
    PT_HIP_CHECK(hipMemcpyAsync(hb, D.bins, P.bins_words * 8, hipMemcpyDeviceToHost, st));
    PT_HIP_CHECK(hipMemcpyAsync(hd, D.out_d, nl * 24, hipMemcpyDeviceToHost, st));
    PT_HIP_CHECK(hipMemcpyAsync(hi, D.out_i, nl * 16, hipMemcpyDeviceToHost, st));
    PT_HIP_CHECK(hipStreamSynchronize(st));
}

void dmumu_merge(const DmumuPlan &P, const unsigned long long *hb, const double *hd, const long long *hi,
                 int64_t *particle_counts, double *particle_means, double *particle_m2s, double *pair_sum_squares,
                 int64_t *pair_counts, int64_t *dmumu_counts, double *dmumu_sum_delta, double *dmumu_sum_delta2) {
    const size_t nb = static_cast<size_t>(P.nbins), nl = static_cast<size_t>(P.nlags);
    for (size_t li = 0; li < nl; ++li) {
        if (P.lags[li] >= P.nsave) continue;  // the reference skips lags beyond the series
        const unsigned long long *bl = hb + li * nb * kBinWords;
        for (size_t b = 0; b < nb; ++b) {
            dmumu_counts[b + nb * li] += static_cast<int64_t>(bl[b]);
            dmumu_sum_delta[b + nb * li] += fix_to_double(bl[nb + 2 * b], bl[nb + 2 * b + 1]);
            dmumu_sum_delta2[b + nb * li] += fix_to_double(bl[3 * nb + 2 * b], bl[3 * nb + 2 * b + 1]);
        }
        combine_welford_host(particle_counts, particle_means, particle_m2s, static_cast<int64_t>(li), hi[li * 2],
                             hd[li * 3], hd[li * 3 + 1]);
        pair_sum_squares[li] += hd[li * 3 + 2];
        pair_counts[li] += hi[li * 2 + 1];
    }
}

// ------------------------------------------ fixed scratch (no allocation) ---
//
// pt_dmumu_accumulate_noalloc places every device work area in this array
// (one copy per GPU, held until the process ends) and its host staging in a
// static host area, so it calls no allocator.  Host code must never pass the
// array's name to a kernel or copy: it is a host-side stand-in.  Its device
// address is looked up once per device (hipGetSymbolAddress) and kept in
// g_noalloc_dev.
constexpr size_t kNoAllocScratchBytes = size_t{32} << 20;
constexpr int64_t kNoAllocMaxLags = 4096;
constexpr int64_t kNoAllocMaxLagBins = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void *g_noalloc_dev[kMaxDevices] = {};
std::mutex g_noalloc_dev_mutex;

struct NoAllocHost {
    unsigned long long bins[kNoAllocMaxLagBins * kBinWords];
    double out_d[kNoAllocMaxLags * 3];
    long long out_i[kNoAllocMaxLags * 2];
};
NoAllocHost g_noalloc_host;
std::mutex g_noalloc_host_mutex;

// Device address of `scratch` on `device` (must be the current device).
void *noalloc_scratch(int device) {
    PT_REQUIRE(device >= 0 && device < kMaxDevices, "pt_dmumu_accumulate_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_dev_mutex);
    if (!g_noalloc_dev[device]) {
        void *ptr = nullptr;
        PT_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PT_REQUIRE(ptr != nullptr, "pt_dmumu_accumulate_noalloc: no device address for the scratch array");
        g_noalloc_dev[device] = ptr;
    }
    return g_noalloc_dev[device];
}

size_t noalloc_bytes(const DmumuPlan &P, bool with_index) {
    return align256(P.work_bytes) + align256(static_cast<size_t>(P.n_sel) * 4) +
           align256(static_cast<size_t>(P.nlags) * 8) + align256(static_cast<size_t>(P.nbins) * 4) +
           (with_index ? align256(static_cast<size_t>(P.n_sel) * 8) : 0);
}

}  // namespace

void noalloc_scratch_init(int device) { noalloc_scratch(device); }

}  // namespace pt

using namespace pt;

extern "C" {

int pt_mu_from_phase(pt_ctx *ctx, const pt_field *field, int64_t chunk, int64_t nsave,
                     int src_dtype, int src_on_device, const void *positions,
                     const void *momenta, int out_dtype, int out_on_device, void *mu_out) {
    return pt_mu_from_phase_layout(ctx, field, chunk, nsave, src_dtype, src_on_device, positions, momenta,
                                   out_dtype, out_on_device, mu_out, 0);
}

int pt_mu_from_phase_layout(pt_ctx *ctx, const pt_field *field, int64_t chunk, int64_t nsave,
                            int src_dtype, int src_on_device, const void *positions,
                            const void *momenta, int out_dtype, int out_on_device, void *mu_out,
                            int sample_major) {
    return guarded([&] {
        PT_REQUIRE(ctx && field && positions && momenta && mu_out, "pt_mu_from_phase: null argument");
        PT_REQUIRE(chunk >= 1 && nsave >= 1, "pt_mu_from_phase: empty chunk");
        PT_REQUIRE(field->storage == PT_STORE_F32_PLANAR || field->storage == PT_STORE_F64_PLANAR,
                   "pt_mu_from_phase: needs a planar field (compute_precision Float32 or Float64)");
        PT_REQUIRE(src_dtype == PT_F32 || src_dtype == PT_F64, "pt_mu_from_phase: bad src dtype");
        PT_REQUIRE(out_dtype == PT_F32 || out_dtype == PT_F64, "pt_mu_from_phase: bad out dtype");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t cnt = static_cast<size_t>(chunk * 3 * nsave);
        DevBuf hp(ctx), hm(ctx), ho(ctx);
        const size_t esz = src_dtype == PT_F64 ? 8 : 4;
        const void *dp = positions, *dm = momenta;
        if (!src_on_device) {
            dp = hp.alloc(cnt * esz);
            dm = hm.alloc(cnt * esz);
            PT_HIP_CHECK(hipMemcpyAsync(const_cast<void *>(dp), positions, cnt * esz, hipMemcpyHostToDevice, st));
            PT_HIP_CHECK(hipMemcpyAsync(const_cast<void *>(dm), momenta, cnt * esz, hipMemcpyHostToDevice, st));
        }
        const size_t osz = out_dtype == PT_F64 ? 8 : 4;
        void *dout = out_on_device ? mu_out : ho.alloc(static_cast<size_t>(chunk * nsave) * osz);
        // samples per thread: long enough that the 24 corner loads at a
        // segment's start (a cache line each) cost less than the segment's
        // own phase-space reads, short enough to leave the chip thousands
        // of waves
        const int32_t seg = 128;
        const int64_t threads = chunk * ((nsave + seg - 1) / seg);
        PT_REQUIRE(threads / 256 < (1LL << 31), "pt_mu_from_phase: chunk too large");
        const unsigned blocks = static_cast<unsigned>((threads + 255) / 256);
        auto run = [&](auto tag_t, auto tag_s, auto tag_o) {
            using T = decltype(tag_t);
            using S = decltype(tag_s);
            using O = decltype(tag_o);
            PlanarGeom<T> g;
            // The reference builds the grid in compute precision T
            // (compute_delta_mu2_curve.jl:55-60, :193-198).
            const T L = static_cast<T>(field->g.L);
            for (int c = 0; c < 3; ++c) {
                g.lo[c] = T(0);
                g.hi[c] = L;
                g.d[c] = static_cast<T>(static_cast<double>(L) / static_cast<double>(field->g.n[c] - 1));
                g.n[c] = field->g.n[c];
            }
            k_mu_from_phase<T, S, O><<<blocks, 256, 0, st>>>(static_cast<const S *>(dp), static_cast<const S *>(dm),
                                                          chunk, nsave, seg, static_cast<const T *>(field->data), g,
                                                          static_cast<O *>(dout), sample_major ? 1 : 0);
        };
        const bool f64 = field->storage == PT_STORE_F64_PLANAR;
        if (f64) {
            if (src_dtype == PT_F64) {
                if (out_dtype == PT_F64) run(double(), double(), double()); else run(double(), double(), float());
            } else {
                if (out_dtype == PT_F64) run(double(), float(), double()); else run(double(), float(), float());
            }
        } else {
            if (src_dtype == PT_F64) {
                if (out_dtype == PT_F64) run(float(), double(), double()); else run(float(), double(), float());
            } else {
                if (out_dtype == PT_F64) run(float(), float(), double()); else run(float(), float(), float());
            }
        }
        PT_HIP_CHECK(hipGetLastError());
        if (!out_on_device)
            PT_HIP_CHECK(hipMemcpyAsync(mu_out, dout, static_cast<size_t>(chunk * nsave) * osz, hipMemcpyDeviceToHost, st));
        PT_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int pt_dmumu_accumulate(pt_ctx *ctx, const float *mu, int mu_on_device, int64_t nsave,
                        int64_t n_sel, int64_t stride_s, int64_t stride_p,
                        const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags,
                        const double *mu_edges, int64_t n_bins, int start_mode, int bin_abs,
                        int64_t *particle_counts, double *particle_means, double *particle_m2s,
                        double *pair_sum_squares, int64_t *pair_counts, int64_t *dmumu_counts,
                        double *dmumu_sum_delta, double *dmumu_sum_delta2) {
    return guarded([&] {
        DmumuPlan P = dmumu_plan(ctx, mu, nsave, n_sel, stride_s, stride_p, particle_index, lag_steps, nlags,
                                 mu_edges, n_bins, start_mode);
        PT_REQUIRE(particle_counts && particle_means && particle_m2s && pair_sum_squares && pair_counts &&
                       dmumu_counts && dmumu_sum_delta && dmumu_sum_delta2,
                   "pt_dmumu_accumulate: null accumulator");
        if (n_sel == 0) return;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf hmu(ctx), hidx(ctx), hlag(ctx), hlo(ctx), hout(ctx), hfin(ctx), htr(ctx);
        DmumuDev D;
        D.mu = to_device(mu, mu_on_device, P.mu_count, hmu, st);
        D.idx = particle_index ? to_device(particle_index, 0, static_cast<size_t>(n_sel), hidx, st) : nullptr;
        // Series stored contiguously per particle (the reference's mu_chunk):
        // one lane per particle would then read a different cache line for
        // every sample, so the selected series are first gathered
        // sample-major (a 4-byte-per-sample copy, against the ~nlags/4 reads
        // of each sample the walk makes).  The walk's per-particle order is
        // unchanged, so the accumulators are the same bit for bit.
        const size_t tr_bytes = static_cast<size_t>(n_sel) * static_cast<size_t>(nsave) * sizeof(float);
        if (stride_s == 1 && nsave > 1 && n_sel > 1 && tr_bytes <= (size_t{4} << 30) && (nsave + 63) / 64 <= 65535) {
            float *tm = static_cast<float *>(htr.alloc(tr_bytes));
            const dim3 grid(static_cast<unsigned>((n_sel + 31) / 32), static_cast<unsigned>((nsave + 63) / 64));
            k_series_sample_major<<<grid, 256, 0, st>>>(D.mu, nsave, n_sel, stride_p, D.idx, tm);
            PT_HIP_CHECK(hipGetLastError());
            P = dmumu_plan(ctx, tm, nsave, n_sel, n_sel, 1, nullptr, lag_steps, nlags, mu_edges, n_bins, start_mode);
            D.mu = tm;
            D.idx = nullptr;
        }
        D.lags = to_device(P.lags, 0, static_cast<size_t>(nlags), hlag, st);
        D.lo = to_device(P.lo, 0, static_cast<size_t>(n_bins), hlo, st);
        dmumu_carve(P, static_cast<char *>(hout.alloc(P.work_bytes)), D);
        D.fin = static_cast<int32_t *>(hfin.alloc(static_cast<size_t>(n_sel) * 4));
        dmumu_enqueue(P, D, bin_abs, start_mode, st);
        const size_t nb = static_cast<size_t>(n_bins), nl = static_cast<size_t>(nlags);
        std::vector<unsigned long long> hb(nl * nb * kBinWords);
        std::vector<double> hd(nl * 3);
        std::vector<long long> hi(nl * 2);
        dmumu_download(P, D, hb.data(), hd.data(), hi.data(), st);
        dmumu_merge(P, hb.data(), hd.data(), hi.data(), particle_counts, particle_means, particle_m2s,
                    pair_sum_squares, pair_counts, dmumu_counts, dmumu_sum_delta, dmumu_sum_delta2);
    });
}

int pt_dmumu_accumulate_noalloc(pt_ctx *ctx, const float *mu, int64_t nsave, int64_t n_sel, int64_t stride_s,
                                int64_t stride_p, const int64_t *particle_index, const int64_t *lag_steps,
                                int64_t nlags, const double *mu_edges, int64_t n_bins, int start_mode, int bin_abs,
                                int64_t *particle_counts, double *particle_means, double *particle_m2s,
                                double *pair_sum_squares, int64_t *pair_counts, int64_t *dmumu_counts,
                                double *dmumu_sum_delta, double *dmumu_sum_delta2) {
    return guarded([&] {
        DmumuPlan P = dmumu_plan(ctx, mu, nsave, n_sel, stride_s, stride_p, particle_index, lag_steps, nlags,
                                 mu_edges, n_bins, start_mode);
        PT_REQUIRE(particle_counts && particle_means && particle_m2s && pair_sum_squares && pair_counts &&
                       dmumu_counts && dmumu_sum_delta && dmumu_sum_delta2,
                   "pt_dmumu_accumulate_noalloc: null accumulator");
        PT_REQUIRE(nlags <= kNoAllocMaxLags && nlags * n_bins <= kNoAllocMaxLagBins,
                   "pt_dmumu_accumulate_noalloc: at most 4096 lags and 65536 (lag, bin) cells");
        const size_t need = noalloc_bytes(P, particle_index != nullptr);
        if (need > kNoAllocScratchBytes)
            throw Error{PT_ERR_MEMORY, "pt_dmumu_accumulate_noalloc: the call needs " + std::to_string(need) +
                                           " bytes of scratch, more than the library's fixed " +
                                           std::to_string(kNoAllocScratchBytes) +
                                           " (fewer particles or lags per call)"};
        if (n_sel == 0) return;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        char *base = static_cast<char *>(noalloc_scratch(ctx->device));
        // carve: [work][fin][lags][lo][index]; every piece 256-byte aligned
        DmumuDev D;
        D.mu = mu;
        dmumu_carve(P, base, D);
        size_t off = align256(P.work_bytes);
        D.fin = reinterpret_cast<int32_t *>(base + off);
        off += align256(static_cast<size_t>(n_sel) * 4);
        int64_t *dl = reinterpret_cast<int64_t *>(base + off);
        off += align256(static_cast<size_t>(nlags) * 8);
        float *dlo = reinterpret_cast<float *>(base + off);
        off += align256(static_cast<size_t>(n_bins) * 4);
        PT_HIP_CHECK(hipMemcpyAsync(dl, P.lags, static_cast<size_t>(nlags) * 8, hipMemcpyHostToDevice, st));
        PT_HIP_CHECK(hipMemcpyAsync(dlo, P.lo, static_cast<size_t>(n_bins) * 4, hipMemcpyHostToDevice, st));
        D.lags = dl;
        D.lo = dlo;
        D.idx = nullptr;
        if (particle_index) {
            int64_t *di = reinterpret_cast<int64_t *>(base + off);
            PT_HIP_CHECK(hipMemcpyAsync(di, particle_index, static_cast<size_t>(n_sel) * 8, hipMemcpyHostToDevice, st));
            D.idx = di;
        }
        // host staging is a fixed static area: calls from several threads serialise here
        std::lock_guard<std::mutex> lock(g_noalloc_host_mutex);
        dmumu_enqueue(P, D, bin_abs, start_mode, st);
        dmumu_download(P, D, g_noalloc_host.bins, g_noalloc_host.out_d, g_noalloc_host.out_i, st);
        dmumu_merge(P, g_noalloc_host.bins, g_noalloc_host.out_d, g_noalloc_host.out_i, particle_counts,
                    particle_means, particle_m2s, pair_sum_squares, pair_counts, dmumu_counts, dmumu_sum_delta,
                    dmumu_sum_delta2);
    });
}

}  // extern "C"

namespace pt {
namespace {
// Src = float momenta [save][3][total] (comps 3, |p| in FP64 first), double
// |p| [save][total] (comps 1), or with kDelta a float relative series
// |p|/p_ref - 1 [save][total] (comps 1).
template <class Src, bool kDelta = false>
int dpp_accumulate(pt_ctx *ctx, const Src *momenta, int on_device, int64_t nsave, int64_t total, int64_t n_sel,
                   const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags, double p0,
                   int64_t *counts, double *sum_delta_p, double *sum_delta_p2, double *sum_delta_p_norm,
                   double *sum_delta_p_norm2, int64_t comps, double p_ref = 0.0) {
    return guarded([&] {
        PT_REQUIRE(ctx && momenta && lag_steps && counts && sum_delta_p && sum_delta_p2 && sum_delta_p_norm &&
                       sum_delta_p_norm2,
                   "pt_dpp_accumulate: null argument");
        PT_REQUIRE(nsave >= 1 && total >= 1 && n_sel >= 0 && nlags >= 1, "pt_dpp_accumulate: empty input");
        PT_REQUIRE(particle_index || n_sel <= total, "pt_dpp_accumulate: n_sel exceeds total");
        std::vector<int64_t> lags(lag_steps, lag_steps + nlags);
        for (int64_t l : lags) PT_REQUIRE(l >= 1, "pt_dpp_accumulate: lag steps must be >= 1");
        int64_t lag_top = 0;
        for (int64_t l : lags) lag_top = std::max(lag_top, std::min(l, nsave));
        if (particle_index)
            for (int64_t j = 0; j < n_sel; ++j)
                PT_REQUIRE(particle_index[j] >= 0 && particle_index[j] < total, "pt_dpp_accumulate: particle index out of range");
        PT_REQUIRE(std::isfinite(p0) && p0 > 0.0, "pt_dpp_accumulate: p0 must be finite and positive");
        if (kDelta) PT_REQUIRE(std::isfinite(p_ref) && p_ref > 0.0, "pt_dpp_accumulate_dprel: p_ref must be finite and positive");
        if (n_sel == 0) return;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf hm(ctx), hidx(ctx), hlag(ctx), hout(ctx), hpm(ctx);
        const Src *dm = to_device(momenta, on_device, static_cast<size_t>(nsave * comps * total), hm, st);
        const int64_t *didx = particle_index ? to_device(particle_index, 0, static_cast<size_t>(n_sel), hidx, st) : nullptr;
        const int64_t *dlag = to_device(lags.data(), 0, lags.size(), hlag, st);
        // FP32 momenta: the walks read an FP64 |p| series
        constexpr bool kMomenta = !kDelta && !std::is_same<Src, double>::value;
        using Series = std::conditional_t<kMomenta, double, Src>;
        // LDS tiles of kp particles when one particle's series fits
        const int kp = dpp_tile_particles(nsave, sizeof(Series));
        // The global walk over FP32 momenta: |p| once per (sample, particle)
        // into an FP64 scratch series, a chunk of particles at a time (at most
        // ~2 GiB).
        int64_t chunk = n_sel;
        if constexpr (kMomenta) {
            const int64_t cap = std::max<int64_t>(kTPB, ((int64_t{1} << 31) / (8 * nsave)) / kTPB * kTPB);
            chunk = std::min(n_sel, cap);
        }
        const int nbx_max = std::max(static_cast<int>((chunk + kTPB - 1) / kTPB), kDppTileBlocks);
        const size_t nl = static_cast<size_t>(nlags), parts = nl * static_cast<size_t>(nbx_max);
        char *base = static_cast<char *>(hout.alloc(parts * (16 + 8) + nl * (16 + 8)));
        auto *part_d = reinterpret_cast<double *>(base);
        auto *part_i = reinterpret_cast<long long *>(part_d + parts * 2);
        auto *out_d = reinterpret_cast<double *>(part_i + parts);
        auto *out_i = reinterpret_cast<long long *>(out_d + nl * 2);
        std::vector<double> hd(nl * 2, 0.0), cd(nl * 2);
        std::vector<long long> hi(nl, 0), ci(nl);
        // merge the partials of one walk in workgroup order and add them up
        auto collect = [&](int nparts) {
            k_merge_dpp<<<static_cast<unsigned>(nlags), kTPB, 0, st>>>(part_d, part_i, static_cast<int>(nlags), nparts,
                                                                       out_d, out_i);
            PT_HIP_CHECK(hipGetLastError());
            PT_HIP_CHECK(hipMemcpyAsync(cd.data(), out_d, nl * 16, hipMemcpyDeviceToHost, st));
            PT_HIP_CHECK(hipMemcpyAsync(ci.data(), out_i, nl * 8, hipMemcpyDeviceToHost, st));
            PT_HIP_CHECK(hipStreamSynchronize(st));
            for (size_t q = 0; q < nl * 2; ++q) hd[q] += cd[q];
            for (size_t q = 0; q < nl; ++q) hi[q] += ci[q];
        };
        // The global walk (series longer than one LDS tile per particle).
        DevBuf hfin(ctx);
        int32_t *fin = nullptr;
        auto run = [&](const auto *pm, int64_t row_len, int64_t nsel_c, const int64_t *idx) {
            using T = std::remove_cv_t<std::remove_pointer_t<decltype(pm)>>;
            const int nbx = static_cast<int>((nsel_c + kTPB - 1) / kTPB);
            const int ngy = static_cast<int>((nlags + kDppLags<T> - 1) / kDppLags<T>);
            if (!fin) fin = static_cast<int32_t *>(hfin.alloc(static_cast<size_t>(std::max(chunk, n_sel)) * 4));
            launch_finite_prefix(pm, nsave, nsel_c, row_len, int64_t{1}, idx, fin, st);
            // the fast path's 32-bit lane offsets (bytes) cover every column
            const int lane32 = row_len < (int64_t{1} << 29) && nsave < (int64_t{1} << 31) ? 1 : 0;
            // buffer loads: the last column's and the largest lag's byte
            // offsets below 2^31
            const bool buf = lane32 && (row_len + lag_top * row_len) * static_cast<int64_t>(sizeof(T)) <
                                           (int64_t{1} << 31);
            auto *kern = buf ? k_dpp<T, kDelta, true> : k_dpp<T, kDelta, false>;
            kern<<<tile_group_blocks(nbx, ngy), kTPB, 0, st>>>(pm, nsave, row_len, nsel_c, idx, dlag,
                                                              static_cast<int>(nlags), part_d, part_i, nbx, ngy, fin,
                                                              lane32);
            PT_HIP_CHECK(hipGetLastError());
            collect(nbx);
        };
        if (kp > 0) {
            // LDS tiles straight from the source (FP32 momenta: |p| while
            // staging): every sample crosses the L2 once
            const int64_t ntiles = (n_sel + kp - 1) / kp;
            const int nparts = static_cast<int>(std::min<int64_t>(ntiles, kDppTileBlocks));
            const size_t lds = static_cast<size_t>(nsave) * kp * sizeof(Series);
            auto launch = [&](auto kernel) {
                kernel<<<static_cast<unsigned>(nparts), kTPB, lds, st>>>(dm, static_cast<int32_t>(nsave), total, n_sel,
                                                                         didx, dlag, static_cast<int>(nlags), part_d,
                                                                         part_i);
            };
            switch (kp) {
                case 16: launch(k_dpp_tile<Series, kDelta, 16, kMomenta>); break;
                case 8: launch(k_dpp_tile<Series, kDelta, 8, kMomenta>); break;
                case 4: launch(k_dpp_tile<Series, kDelta, 4, kMomenta>); break;
                case 2: launch(k_dpp_tile<Series, kDelta, 2, kMomenta>); break;
                default: launch(k_dpp_tile<Series, kDelta, 1, kMomenta>); break;
            }
            PT_HIP_CHECK(hipGetLastError());
            collect(nparts);
        } else if constexpr (!kMomenta) {
            run(dm, total, n_sel, didx);
        } else {
            auto *pm = static_cast<double *>(hpm.alloc(static_cast<size_t>(nsave * chunk) * 8));
            for (int64_t c0 = 0; c0 < n_sel; c0 += chunk) {
                const int64_t nc = std::min(chunk, n_sel - c0);
                const unsigned blocks = static_cast<unsigned>((nsave * nc + 255) / 256);
                k_pmag<<<blocks, 256, 0, st>>>(dm, nsave, total, c0, nc, didx, pm);
                PT_HIP_CHECK(hipGetLastError());
                run(pm, nc, nc, nullptr);
            }
        }
        const double rn = p_ref / p0;  // delta units -> Delta p / p0
        for (size_t li = 0; li < nl; ++li) {
            if (lags[li] >= nsave) continue;
            counts[li] += hi[li];
            if constexpr (kDelta) {
                sum_delta_p[li] += p_ref * hd[li * 2 + 0];
                sum_delta_p2[li] += p_ref * p_ref * hd[li * 2 + 1];
                sum_delta_p_norm[li] += rn * hd[li * 2 + 0];
                sum_delta_p_norm2[li] += rn * rn * hd[li * 2 + 1];
            } else {
                sum_delta_p[li] += hd[li * 2 + 0];
                sum_delta_p2[li] += hd[li * 2 + 1];
                sum_delta_p_norm[li] += hd[li * 2 + 0] / p0;
                sum_delta_p_norm2[li] += hd[li * 2 + 1] / p0 / p0;
            }
        }
    });
}

}  // namespace
}  // namespace pt

extern "C" {

int pt_dpp_accumulate(pt_ctx *ctx, const float *momenta, int on_device, int64_t nsave,
                      int64_t total, int64_t n_sel, const int64_t *particle_index,
                      const int64_t *lag_steps, int64_t nlags, double p0, int64_t *counts,
                      double *sum_delta_p, double *sum_delta_p2, double *sum_delta_p_norm,
                      double *sum_delta_p_norm2) {
    return pt::dpp_accumulate(ctx, momenta, on_device, nsave, total, n_sel, particle_index, lag_steps, nlags, p0,
                              counts, sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2, 3);
}

int pt_dpp_accumulate_pmag(pt_ctx *ctx, const double *pmag, int on_device, int64_t nsave,
                           int64_t total, int64_t n_sel, const int64_t *particle_index,
                           const int64_t *lag_steps, int64_t nlags, double p0, int64_t *counts,
                           double *sum_delta_p, double *sum_delta_p2, double *sum_delta_p_norm,
                           double *sum_delta_p_norm2) {
    return pt::dpp_accumulate(ctx, pmag, on_device, nsave, total, n_sel, particle_index, lag_steps, nlags, p0,
                              counts, sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2, 1);
}

int pt_dpp_accumulate_dprel(pt_ctx *ctx, const float *dprel, int on_device, int64_t nsave, int64_t total,
                            int64_t n_sel, const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags,
                            double p_ref, double p0, int64_t *counts, double *sum_delta_p, double *sum_delta_p2,
                            double *sum_delta_p_norm, double *sum_delta_p_norm2) {
    return pt::dpp_accumulate<float, true>(ctx, dprel, on_device, nsave, total, n_sel, particle_index, lag_steps,
                                           nlags, p0, counts, sum_delta_p, sum_delta_p2, sum_delta_p_norm,
                                           sum_delta_p_norm2, 1, p_ref);
}

int pt_energy_snapshots(pt_ctx *ctx, const float *momenta, int on_device, int64_t nsave,
                        int64_t total, int64_t n_sel, const int64_t *particle_index,
                        const int64_t *snapshot_idx, int64_t k, double *out, int64_t col0) {
    return guarded([&] {
        PT_REQUIRE(ctx && momenta && snapshot_idx && out, "pt_energy_snapshots: null argument");
        PT_REQUIRE(k >= 1 && n_sel >= 0 && col0 >= 0, "pt_energy_snapshots: bad sizes");
        for (int64_t q = 0; q < k; ++q)
            PT_REQUIRE(snapshot_idx[q] >= 0 && snapshot_idx[q] < nsave, "pt_energy_snapshots: snapshot index out of range");
        if (particle_index)
            for (int64_t j = 0; j < n_sel; ++j)
                PT_REQUIRE(particle_index[j] >= 0 && particle_index[j] < total, "pt_energy_snapshots: particle index out of range");
        else
            PT_REQUIRE(n_sel <= total, "pt_energy_snapshots: n_sel exceeds total");
        if (n_sel == 0) return;
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf hm(ctx), hidx(ctx), hs(ctx), ho(ctx);
        const float *dm = to_device(momenta, on_device, static_cast<size_t>(nsave * 3 * total), hm, st);
        const int64_t *didx = particle_index ? to_device(particle_index, 0, static_cast<size_t>(n_sel), hidx, st) : nullptr;
        const int64_t *dsn = to_device(snapshot_idx, 0, static_cast<size_t>(k), hs, st);
        double *dout = static_cast<double *>(ho.alloc(static_cast<size_t>(n_sel * k) * 8));
        const unsigned blocks = static_cast<unsigned>((n_sel * k + 255) / 256);
        k_energy<<<blocks, 256, 0, st>>>(dm, total, n_sel, didx, dsn, k, dout);
        PT_HIP_CHECK(hipGetLastError());
        PT_HIP_CHECK(hipMemcpyAsync(out + k * col0, dout, static_cast<size_t>(n_sel * k) * 8, hipMemcpyDeviceToHost, st));
        PT_HIP_CHECK(hipStreamSynchronize(st));
    });
}

}  // extern "C"
