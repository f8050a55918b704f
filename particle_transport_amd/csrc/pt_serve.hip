// Continuous-batching campaign server.
//
// Requests (field x energy x injection x duration, uneven particle counts)
// are admitted FIFO against an HBM budget for their particle state plus
// cadence cache, and every admitted request is advanced by one fused launch
// per tick: the grid is the concatenation of every request's 256-particle
// blocks, each block finds its request in a device table (binary search on
// the block prefix) and runs the production step loop (core::run_lane) with
// that request's own field, dt, charge, mass, duration and cadence.  Up to
// two ticks are in flight: pt_serve_step completes the older one when two are
// enqueued (retiring the requests whose last steps it carried), admits queued
// requests and enqueues the next tick, so the device always has work while
// the host runs; a step with nothing new to enqueue completes everything in
// flight.  Requests arriving while others run join at the next tick.
//
// The tick launch is an ordinary kernel rather than a spin-waiting resident
// kernel: the host changes the request mix between ticks, and no kernel ever
// waits on host memory.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "pt_push_core.hpp"

namespace pt {
namespace {

using namespace core;

struct DevReq {
    int32_t *cell[3];
    float *off[3];
    float *comp[3];         // compensation of the offset sums (SplitState::comp)
    float *U[3];            // momentum / (m c) of this request's species
    double *P2;             // |U|^2 in FP64
    uint8_t *alive;
    unsigned long long *deaths;
    int *bad;               // set by the injection kernel on zero/non-finite local B
    void *out;              // cadence cache [save][n]
    int64_t n, total, stride, block0;
    Cadence cad;
    int32_t out_kind, periodic;
    double p_ref;  // PT_OUT_DPREL: delta = |p| / p_ref - 1
    FieldView fv;
    SplitGeom g;
    Phys ph;
};

// CUBIC: every request of the launch has cubic cells (the same c dt / dx on
// every axis), so the position step factor is formed once per push.
template <int STORE, bool CUBIC>
__device__ __forceinline__ void serve_tick(const DevReq *__restrict__ reqs, int nreq) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = nreq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].block0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const DevReq &R = reqs[lo];
    const int64_t p = (b - R.block0) * blockDim.x + threadIdx.x;
    const int64_t n = R.n;
    // lanes past the request's last particle stay as dead helpers of the
    // wave-cooperative reload; a wave without any particle leaves at once
    const bool valid = p < n;
    if (__ballot(valid) == 0) return;
    __shared__ ReloadLds lds[4];
    const int64_t q = valid ? p : 0;
    bool alive = valid && R.alive[q] != 0;
    int32_t cell[3];
    float off[3], comp[3], U[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cell[c] = R.cell[c][q];
        off[c] = R.off[c][q];
        comp[c] = R.comp[c][q];
        U[c] = R.U[c][q];
    }
    double P2 = R.P2[q];
    const int kind = R.out_kind;
    void *out = R.out;
    const float qsign = R.ph.qsign;
    const double pout = R.ph.pout;
    const double p_ref = R.p_ref;
    // The request's parameters as values: read through R, the loop's stores
    // (cache, death counter) could alias the table and force reloads each push.
    const FieldView fv = R.fv;
    const SplitGeom g = R.g;
    const Phys ph = R.ph;
    const Cadence cad = R.cad;
    const int periodic = R.periodic;
    unsigned long long *const deaths = R.deaths;
    run_lane<STORE, CUBIC>(
        cell, off, comp, U, P2, alive, fv, g, ph, cad, periodic, &lds[threadIdx.x >> 6],
        [&](int64_t slot, bool alv, const int32_t *, const float *, const float *, const float *u, double p2,
            const float *v) {
            if (!valid) return;
            if (kind == PT_OUT_MU) {
                static_cast<float *>(out)[slot * n + p] = alv ? mu_f32(u[0], u[1], u[2], v[0], v[1], v[2], qsign) : NAN;
            } else if (kind == PT_OUT_DPREL) {
                static_cast<float *>(out)[slot * n + p] =
                    alv ? static_cast<float>(sqrt(p2) * pout / p_ref - 1.0) : NAN;
            } else {
                static_cast<double *>(out)[slot * n + p] = alv ? sqrt(p2) * pout : NAN;
            }
        },
        [&](int64_t) { atomicAdd(deaths, 1ull); });
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R.cell[c][p] = cell[c];
        R.off[c][p] = off[c];
        R.comp[c][p] = comp[c];
        R.U[c][p] = U[c];
    }
    R.P2[p] = P2;
    R.alive[p] = alive ? 1 : 0;
}

// One tick over all running requests (waves per SIMD: kPushWaves).
template <int STORE, bool CUBIC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPushWaves<STORE>, 8)))
k_serve(const DevReq *__restrict__ reqs, int nreq) {
    serve_tick<STORE, CUBIC>(reqs, nreq);
}

// Initial state of one request: split position from metres; momentum either
// given (isotropic or any host-built state) or built on the device for a
// fixed pitch cosine mu0 about the local field (fixed_mu_unit_vector,
// phase_space_gpu_runner.jl:170-194) from the request's gyrophase draws.
template <int STORE>
__global__ void k_serve_init(DevReq R, const double *__restrict__ init, int fixed_mu, float mu0, double pnorm,
                             double p_magnitude) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = R.n;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        int32_t cell;
        float off, comp;
        split_coord(init[c * n + i] / R.g.d[c], R.g.nc[c], R.periodic != 0, cell, off, comp);
        R.cell[c][i] = cell;
        R.off[c][i] = off;
        R.comp[c][i] = comp;
    }
    if (!fixed_mu) {
        double p2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double pc = init[(3 + c) * n + i] / pnorm;  // as pt_ens_create converts
            p2 += pc * pc;
            R.U[c][i] = static_cast<float>(pc);
        }
        R.P2[i] = p2;
        return;
    }
    int32_t g[3];
    float t[3];
    for (int c = 0; c < 3; ++c) gather_coord(R.cell[c][i], R.off[c][i], R.g.nc[c], g[c], t[c]);
    float sv[6][8], a[8], b[3];
    load_corners<STORE>(R.fv, g[0], g[1], g[2], sv);
    for (int c = 0; c < 3; ++c) {
        cell_coeffs<STORE>(R.fv, g[0], g[1], g[2], c, sv[c], R.fv.scale[c], a);
        b[c] = eval_poly(a, t[0], t[1], t[2]);
    }
    const float bn = sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (!isfinite(bn) || bn == 0.0f) {
        atomicExch(R.bad, 1);
        for (int c = 0; c < 3; ++c) R.U[c][i] = 0.0f;
        R.P2[i] = 0.0;
        return;
    }
    const float bhx = b[0] / bn, bhy = b[1] / bn, bhz = b[2] / bn;
    float e1x, e1y, e1z;
    if (fabsf(bhx) < 0.9f) {
        e1x = 0.0f;
        e1y = bhz;
        e1z = -bhy;
    } else {
        e1x = -bhz;
        e1y = 0.0f;
        e1z = bhx;
    }
    const float e1n = sqrtf(e1x * e1x + e1y * e1y + e1z * e1z);
    e1x /= e1n;
    e1y /= e1n;
    e1z /= e1n;
    const float e2x = bhy * e1z - bhz * e1y;
    const float e2y = bhz * e1x - bhx * e1z;
    const float e2z = bhx * e1y - bhy * e1x;
    const float phi = static_cast<float>(init[3 * n + i]);
    const float perp = sqrtf(fmaxf(0.0f, 1.0f - mu0 * mu0));
    const float cp = cosf(phi), sp = sinf(phi);
    const double pm = p_magnitude / pnorm;
    R.P2[i] = pm * pm;
    const float scale = static_cast<float>(pm);
    R.U[0][i] = scale * (mu0 * bhx + perp * (cp * e1x + sp * e2x));
    R.U[1][i] = scale * (mu0 * bhy + perp * (cp * e1y + sp * e2y));
    R.U[2][i] = scale * (mu0 * bhz + perp * (cp * e1z + sp * e2z));
}

template <class F>
int guarded(F &&fn) {
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

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Best-fit allocator over the server's one device arena (offsets in bytes,
// multiples of kAlign).  Free blocks are kept by offset, to merge a released
// block with its neighbours, and by length, to place a request in the
// smallest block that holds it.
class Arena {
  public:
    void reset(size_t bytes) {
        by_off_.clear();
        by_len_.clear();
        if (bytes) insert(0, bytes);
    }
    bool alloc(size_t n, size_t &off) {
        auto it = by_len_.lower_bound(n);
        if (it == by_len_.end()) return false;
        const size_t len = it->first;
        off = it->second;
        erase(off, len);
        if (len > n) insert(off + n, len - n);
        return true;
    }
    void release(size_t off, size_t n) {
        auto next = by_off_.lower_bound(off);
        if (next != by_off_.end() && off + n == next->first) {
            const size_t o = next->first, l = next->second;
            erase(o, l);
            n += l;
        }
        auto after = by_off_.lower_bound(off);
        if (after != by_off_.begin()) {
            auto prev = std::prev(after);
            if (prev->first + prev->second == off) {
                const size_t o = prev->first, l = prev->second;
                erase(o, l);
                off = o;
                n += l;
            }
        }
        insert(off, n);
    }

  private:
    void insert(size_t off, size_t len) {
        by_off_.emplace(off, len);
        by_len_.emplace(len, off);
    }
    void erase(size_t off, size_t len) {
        by_off_.erase(off);
        auto range = by_len_.equal_range(len);
        for (auto it = range.first; it != range.second; ++it)
            if (it->second == off) {
                by_len_.erase(it);
                break;
            }
    }
    std::map<size_t, size_t> by_off_;
    std::multimap<size_t, size_t> by_len_;
};

// One launch of k_serve: its request table (pinned host copy and device
// copy), events and the work it carries (request id, first step, steps).
struct Tick {
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevReq *d_table = nullptr, *h_table = nullptr;
    int cap = 0;
    std::vector<std::tuple<int64_t, int64_t, int64_t>> work;
};

struct SReq {
    int64_t id = 0;
    int state = PT_REQ_QUEUED;
    pt_serve_request d{};
    std::vector<double> init;
    char *mem = nullptr;
    size_t bytes = 0;
    int64_t si = 1;          // next step not yet completed on the device
    int64_t si_planned = 1;  // next step not yet enqueued
    int64_t nsave = 0;
    int64_t alive_final = 0;
    bool failed = false;
    DevReq dev{};
};

}  // namespace
}  // namespace pt

struct pt_serve {
    pt_ctx *ctx = nullptr;
    int64_t budget = 0;
    int32_t steps_per_tick = 200;
    int32_t max_active = 1024;
    std::vector<const pt_field *> fields;
    int storage = -1;
    std::map<int64_t, pt::SReq> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    int64_t reserved = 0, next_id = 1, n_done = 0;
    std::vector<int64_t> newly_done;  // DONE since the last pt_serve_take_done
    // Two ticks in flight: step() enqueues the next tick before waiting for
    // the previous one, so the device always has a tick queued while the
    // host admits, completes and hands out requests.
    pt::Tick tick[2];
    std::deque<int> inflight;  // slots of enqueued, not yet completed ticks, oldest first
    int64_t particle_steps = 0;
    double kernel_ms = 0.0;
    // Requests' memory comes from one device arena reserved when the server
    // is created (the budget plus kArenaSlack for fragmentation), placed by
    // best fit.  Admission and release are host bookkeeping: no allocation,
    // free or device-wide synchronisation while the server runs, so a tick
    // never waits for work on other streams (a statistics product, say).
    char *arena_base = nullptr;
    size_t arena_bytes = 0;
    pt::Arena arena;
};

namespace pt {
namespace {

size_t request_bytes(const pt_serve_request &d, int64_t nsave) {
    const size_t n = static_cast<size_t>(d.n);
    const size_t esz = d.out_kind == PT_OUT_PMAG ? 8 : 4;
    size_t b = align_up(n * 8) + 12 * align_up(n * 4) + align_up(n) + kAlign;
    if (d.out_kind != PT_OUT_NONE) b += align_up(static_cast<size_t>(nsave) * n * esz);
    return b;
}

void fill_dev(pt_serve *srv, SReq &r) {
    const pt_field *f = srv->fields[static_cast<size_t>(r.d.field_id)];
    const size_t n = static_cast<size_t>(r.d.n);
    char *m = r.mem;
    DevReq &D = r.dev;
    D.P2 = reinterpret_cast<double *>(m);
    m += align_up(n * 8);
    for (int c = 0; c < 3; ++c) {
        D.cell[c] = reinterpret_cast<int32_t *>(m);
        m += align_up(n * 4);
    }
    for (int c = 0; c < 3; ++c) {
        D.off[c] = reinterpret_cast<float *>(m);
        m += align_up(n * 4);
    }
    for (int c = 0; c < 3; ++c) {
        D.U[c] = reinterpret_cast<float *>(m);
        m += align_up(n * 4);
    }
    for (int c = 0; c < 3; ++c) {
        D.comp[c] = reinterpret_cast<float *>(m);
        m += align_up(n * 4);
    }
    D.alive = reinterpret_cast<uint8_t *>(m);
    m += align_up(n);
    D.deaths = reinterpret_cast<unsigned long long *>(m);
    D.bad = reinterpret_cast<int *>(m + 8);
    m += kAlign;
    D.out = r.d.out_kind != PT_OUT_NONE ? m : nullptr;
    D.n = r.d.n;
    D.total = r.d.total_steps;
    D.stride = r.d.save_stride;
    D.out_kind = r.d.out_kind;
    D.p_ref = r.d.out_p_ref;
    D.periodic = r.d.periodic ? 1 : 0;
    D.fv = make_field_view(f);
    for (int c = 0; c < 3; ++c) {
        D.g.nc[c] = static_cast<int32_t>(f->g.n[c] - 1);
        D.g.k[c] = static_cast<float>(kC * r.d.dt_s / f->g.d[c]);
        D.g.d[c] = f->g.d[c];
    }
    const double bfac = r.d.charge_C * r.d.dt_s / (2.0 * r.d.mass_kg);
    for (int c = 0; c < 6; ++c) {
        D.ph.fac[c] = static_cast<float>(c < 3 ? bfac : 1.0 / kC);
    }
    D.ph.qsign = r.d.charge_C > 0 ? 1.0f : -1.0f;
    D.ph.mr = 1.0f;  // state is already in units of this species' m c
    D.ph.pout = r.d.mass_kg * kC;
}

template <int STORE>
void launch_init(pt_serve *srv, SReq &r, const double *dinit, int fixed_mu) {
    const unsigned blocks = static_cast<unsigned>((r.d.n + 255) / 256);
    k_serve_init<STORE><<<blocks, 256, 0, srv->ctx->stream>>>(r.dev, dinit, fixed_mu, static_cast<float>(r.d.mu0),
                                                               r.d.mass_kg * kC, r.d.p_magnitude);
    PT_HIP_CHECK(hipGetLastError());
}

// The arena is the budget plus this fraction of it, so that best-fit
// placement of requests of every size still finds room for the budget.
constexpr double kArenaSlack = 0.125;

// Places the request in the arena (false: no free block holds it yet).
bool admit(pt_serve *srv, SReq &r) {
    hipStream_t st = srv->ctx->stream;
    size_t off = 0;
    if (!srv->arena.alloc(r.bytes, off)) return false;
    r.mem = srv->arena_base + off;
    fill_dev(srv, r);
    PT_HIP_CHECK(hipMemsetAsync(r.dev.alive, 1, static_cast<size_t>(r.d.n), st));
    PT_HIP_CHECK(hipMemsetAsync(r.dev.deaths, 0, kAlign, st));  // deaths and the injection flag
    double *dinit;
    int fixed_mu = r.d.inject_fixed_mu;
    if (r.d.inject_on_device) {
        // the initial state as pt_inject draws it, then converted like explicit
        // host momenta; a fixed-mu failure is read when the request completes
        dinit = static_cast<double *>(ctx_scratch(srv->ctx, static_cast<size_t>(r.d.n) * 6 * sizeof(double)));
        launch_inject(srv->fields[static_cast<size_t>(r.d.field_id)], r.d.inject, r.d.inject_first_index, r.d.n,
                      dinit, r.dev.bad, st);
        fixed_mu = 0;
    } else {
        dinit = static_cast<double *>(ctx_scratch(srv->ctx, r.init.size() * sizeof(double)));
        PT_HIP_CHECK(hipMemcpyAsync(dinit, r.init.data(), r.init.size() * sizeof(double), hipMemcpyHostToDevice,
                                    st));
    }
    switch (srv->storage) {
        case PT_STORE_F32_BRICK: launch_init<PT_STORE_F32_BRICK>(srv, r, dinit, fixed_mu); break;
        case PT_STORE_BF16_BRICK: launch_init<PT_STORE_BF16_BRICK>(srv, r, dinit, fixed_mu); break;
        case PT_STORE_F16_BRICK: launch_init<PT_STORE_F16_BRICK>(srv, r, dinit, fixed_mu); break;
        default: throw Error{PT_ERR_STATE, "server storage not set"};
    }
    if (fixed_mu) {
        int bad = 0;
        PT_HIP_CHECK(hipMemcpyAsync(&bad, r.dev.bad, sizeof(int), hipMemcpyDeviceToHost, st));
        PT_HIP_CHECK(hipStreamSynchronize(st));
        if (bad) r.failed = true;
    }
    std::vector<double>().swap(r.init);
    srv->reserved += static_cast<int64_t>(r.bytes);
    r.state = PT_REQ_RUNNING;
    return true;
}

// Wait for the oldest in-flight tick and retire its work: requests whose
// last step it carried are done.
void complete_oldest(pt_serve *srv) {
    if (srv->inflight.empty()) return;
    Tick &t = srv->tick[srv->inflight.front()];
    PT_HIP_CHECK(hipEventSynchronize(t.ev1));
    float ms = 0.f;
    PT_HIP_CHECK(hipEventElapsedTime(&ms, t.ev0, t.ev1));
    srv->kernel_ms += ms;
    srv->inflight.pop_front();
    for (const auto &w : t.work) {
        SReq &r = srv->reqs.at(std::get<0>(w));
        const int64_t si0 = std::get<1>(w), K = std::get<2>(w);
        const int64_t pushes = std::max<int64_t>(0, std::min(K, r.d.total_steps - si0));
        srv->particle_steps += pushes * r.d.n;
        r.si = si0 + K;
        if (r.si > r.d.total_steps) {
            unsigned long long flags[2] = {0, 0};  // deaths, then the injection flag (its int in the low half)
            PT_HIP_CHECK(hipMemcpy(flags, r.dev.deaths, 16, hipMemcpyDeviceToHost));
            const unsigned long long deaths = flags[0];
            if (static_cast<uint32_t>(flags[1]) != 0) r.failed = true;
            r.alive_final = r.d.n - static_cast<int64_t>(deaths);
            r.state = PT_REQ_DONE;
            srv->n_done += 1;
            srv->newly_done.push_back(r.id);
            srv->active.erase(std::remove(srv->active.begin(), srv->active.end(), r.id), srv->active.end());
        }
    }
    t.work.clear();
}

void complete_all(pt_serve *srv) {
    while (!srv->inflight.empty()) complete_oldest(srv);
}

template <int STORE>
void launch_tick(pt_serve *srv, const DevReq *table, int nreq, int64_t blocks, bool cubic) {
    const unsigned grid = static_cast<unsigned>(blocks);
    if (cubic)
        k_serve<STORE, true><<<grid, 256, 0, srv->ctx->stream>>>(table, nreq);
    else
        k_serve<STORE, false><<<grid, 256, 0, srv->ctx->stream>>>(table, nreq);
    PT_HIP_CHECK(hipGetLastError());
}

// Enqueue one tick for every running request with steps not yet enqueued,
// in a free slot (the caller has completed its previous tick).
void enqueue_tick(pt_serve *srv) {
    std::vector<SReq *> work;
    for (int64_t id : srv->active) {
        SReq &r = srv->reqs.at(id);
        if (r.si_planned <= r.d.total_steps) work.push_back(&r);
    }
    const int nreq = static_cast<int>(work.size());
    if (nreq == 0) return;
    // Longest blocks first: blocks start in grid order, so the slow ones
    // (high energy: a wave with lanes changing cell runs the cooperative
    // reload on most pushes, up to three times the cost of a push without
    // it) go first and the cheap ones fill the chip as the launch drains.
    // The cost model per block and step is 1 + 2.5 P, with P = 1 - (1 -
    // 1.5 k)^64 the chance that one of a wave's 64 particles changes cell in
    // a push (k = c dt / dx cells per push at v ~ c; 1.5 = the mean sum of
    // |v_c| / v over axes for isotropic directions).  Where a block runs in
    // the grid changes no result.
    auto block_cost = [&](const SReq *r) {
        const double k = std::min(1.0, 1.5 * static_cast<double>(r->dev.g.k[0]));
        const double p = 1.0 - std::pow(1.0 - k, 64.0);
        const int64_t K = std::min<int64_t>(srv->steps_per_tick, r->d.total_steps - r->si_planned + 1);
        return static_cast<double>(K) * (1.0 + 2.5 * p);
    };
    std::stable_sort(work.begin(), work.end(),
                     [&](const SReq *a, const SReq *b) { return block_cost(a) > block_cost(b); });
    const int slot = srv->inflight.empty() ? 0 : 1 - srv->inflight.back();
    Tick &t = srv->tick[slot];
    if (t.cap < nreq) {
        if (t.d_table) (void)hipFree(t.d_table);
        if (t.h_table) (void)hipHostFree(t.h_table);
        t.d_table = nullptr;
        t.h_table = nullptr;
        const int cap = std::max(nreq, 2 * t.cap);
        PT_HIP_CHECK(hipMalloc(&t.d_table, sizeof(DevReq) * cap));
        PT_HIP_CHECK(hipHostMalloc(&t.h_table, sizeof(DevReq) * cap, hipHostMallocDefault));
        t.cap = cap;
    }
    int64_t blocks = 0;
    bool cubic = true;
    int with_cells = 0;
    t.work.clear();
    for (int i = 0; i < nreq; ++i) {
        SReq &r = *work[static_cast<size_t>(i)];
        const float *k = r.dev.g.k;
        cubic = cubic && k[0] == k[1] && k[0] == k[2];
        with_cells += r.dev.fv.cells ? 1 : 0;
        const int64_t K = std::min<int64_t>(srv->steps_per_tick, r.d.total_steps - r.si_planned + 1);
        r.dev.cad = make_cadence(r.si_planned, static_cast<int>(K), r.d.total_steps, r.d.save_stride,
                                 r.d.out_kind != PT_OUT_NONE, 0);
        r.dev.block0 = blocks;
        blocks += (r.d.n + 255) / 256;
        t.h_table[i] = r.dev;
        t.work.emplace_back(r.id, r.si_planned, K);
        r.si_planned += K;
    }
    PT_REQUIRE(blocks < (1LL << 31), "pt_serve_step: tick grid too large");
    hipStream_t st = srv->ctx->stream;
    PT_HIP_CHECK(hipEventRecord(t.ev0, st));
    PT_HIP_CHECK(hipMemcpyAsync(t.d_table, t.h_table, sizeof(DevReq) * nreq, hipMemcpyHostToDevice, st));
    // One kernel serves every request of the tick: all of them read cell
    // tables or none does (a table built or freed while its field serves
    // requests is refused here).
    PT_REQUIRE(with_cells == 0 || with_cells == nreq,
               "pt_serve_step: the running requests' fields differ in having a cell table");
    if (with_cells) {
        launch_tick<kStoreCells>(srv, t.d_table, nreq, blocks, cubic);
        PT_HIP_CHECK(hipEventRecord(t.ev1, st));
        srv->inflight.push_back(slot);
        return;
    }
    switch (srv->storage) {
        case PT_STORE_F32_BRICK: launch_tick<PT_STORE_F32_BRICK>(srv, t.d_table, nreq, blocks, cubic); break;
        case PT_STORE_BF16_BRICK: launch_tick<PT_STORE_BF16_BRICK>(srv, t.d_table, nreq, blocks, cubic); break;
        case PT_STORE_F16_BRICK: launch_tick<PT_STORE_F16_BRICK>(srv, t.d_table, nreq, blocks, cubic); break;
        default: throw Error{PT_ERR_STATE, "server storage not set"};
    }
    PT_HIP_CHECK(hipEventRecord(t.ev1, st));
    srv->inflight.push_back(slot);
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" {

int pt_serve_create(pt_ctx *ctx, int64_t hbm_budget_bytes, int32_t steps_per_tick, int32_t max_active,
                    pt_serve **out) {
    return guarded([&] {
        PT_REQUIRE(ctx && out, "pt_serve_create: null argument");
        PT_REQUIRE(hbm_budget_bytes > 0, "pt_serve_create: budget must be positive");
        PT_REQUIRE(steps_per_tick >= 1, "pt_serve_create: steps_per_tick must be >= 1");
        PT_REQUIRE(max_active >= 1, "pt_serve_create: max_active must be >= 1");
        PT_HIP_CHECK(hipSetDevice(ctx->device));
        auto *s = new pt_serve();
        s->ctx = ctx;
        s->budget = hbm_budget_bytes;
        s->steps_per_tick = steps_per_tick;
        s->max_active = max_active;
        for (auto &t : s->tick) {
            PT_HIP_CHECK(hipEventCreate(&t.ev0));
            PT_HIP_CHECK(hipEventCreate(&t.ev1));
        }
        s->arena_bytes = align_up(static_cast<size_t>(static_cast<double>(hbm_budget_bytes) * (1.0 + kArenaSlack)));
        void *mem = nullptr;
        const hipError_t e = hipMalloc(&mem, s->arena_bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            for (auto &t : s->tick) {
                (void)hipEventDestroy(t.ev0);
                (void)hipEventDestroy(t.ev1);
            }
            const size_t want = s->arena_bytes;
            delete s;
            throw Error{PT_ERR_MEMORY, "pt_serve_create: cannot reserve the " + std::to_string(want) +
                                           "-byte arena for the HBM budget: " + hipGetErrorString(e)};
        }
        s->arena_base = static_cast<char *>(mem);
        s->arena.reset(s->arena_bytes);
        ctx->bytes_in_use += static_cast<int64_t>(s->arena_bytes);
        *out = s;
    });
}

int pt_serve_destroy(pt_serve *srv) {
    return guarded([&] {
        if (!srv) return;
        try {
            complete_all(srv);
        } catch (...) {
        }
        (void)hipStreamSynchronize(srv->ctx->stream);
        if (srv->arena_base) {
            (void)hipFree(srv->arena_base);
            srv->ctx->bytes_in_use -= static_cast<int64_t>(srv->arena_bytes);
        }
        for (auto &t : srv->tick) {
            if (t.d_table) (void)hipFree(t.d_table);
            if (t.h_table) (void)hipHostFree(t.h_table);
            if (t.ev0) (void)hipEventDestroy(t.ev0);
            if (t.ev1) (void)hipEventDestroy(t.ev1);
        }
        delete srv;
    });
}

int pt_serve_add_field(pt_serve *srv, const pt_field *field, int32_t *field_id) {
    return guarded([&] {
        PT_REQUIRE(srv && field && field_id, "pt_serve_add_field: null argument");
        PT_REQUIRE(field->storage == PT_STORE_BF16_BRICK || field->storage == PT_STORE_F16_BRICK ||
                       field->storage == PT_STORE_F32_BRICK,
                   "pt_serve_add_field: serving needs a BRICK field");
        PT_REQUIRE(srv->storage < 0 || srv->storage == field->storage,
                   "pt_serve_add_field: all fields of a server share one storage kind");
        PT_REQUIRE(srv->fields.empty() || (srv->fields.front()->cells != nullptr) == (field->cells != nullptr),
                   "pt_serve_add_field: all fields of a server have a cell table or none does");
        PT_REQUIRE(field->ctx == srv->ctx, "pt_serve_add_field: field belongs to another context");
        srv->storage = field->storage;
        srv->fields.push_back(field);
        *field_id = static_cast<int32_t>(srv->fields.size() - 1);
    });
}

int pt_serve_submit(pt_serve *srv, const pt_serve_request *req, int64_t *request_id) {
    return guarded([&] {
        PT_REQUIRE(srv && req && request_id, "pt_serve_submit: null argument");
        const pt_serve_request &d = *req;
        PT_REQUIRE(d.field_id >= 0 && d.field_id < static_cast<int32_t>(srv->fields.size()), "pt_serve_submit: unknown field id");
        PT_REQUIRE(d.n >= 1 && d.n < (1LL << 31), "pt_serve_submit: particle count out of range");
        PT_REQUIRE(std::isfinite(d.dt_s) && d.dt_s > 0, "pt_serve_submit: dt must be finite and positive");
        PT_REQUIRE(std::isfinite(d.mass_kg) && d.mass_kg > 0, "pt_serve_submit: mass must be positive");
        PT_REQUIRE(std::isfinite(d.charge_C) && d.charge_C != 0, "pt_serve_submit: charge must be non-zero");
        PT_REQUIRE(d.total_steps >= 1 && d.save_stride >= 1, "pt_serve_submit: total_steps and save_stride must be >= 1");
        PT_REQUIRE(d.out_kind == PT_OUT_NONE || d.out_kind == PT_OUT_MU || d.out_kind == PT_OUT_PMAG ||
                       d.out_kind == PT_OUT_DPREL,
                   "pt_serve_submit: out_kind must be PT_OUT_NONE, PT_OUT_MU, PT_OUT_PMAG or PT_OUT_DPREL");
        if (d.out_kind == PT_OUT_DPREL)
            PT_REQUIRE(std::isfinite(d.out_p_ref) && d.out_p_ref > 0, "pt_serve_submit: out_p_ref must be positive");
        PT_REQUIRE(d.inject_on_device == 0 || d.inject_on_device == 1, "pt_serve_submit: inject_on_device must be 0 or 1");
        const pt_field *f = srv->fields[static_cast<size_t>(d.field_id)];
        if (d.inject_on_device) {
            PT_REQUIRE(d.inject_first_index >= 0, "pt_serve_submit: inject_first_index must be >= 0");
            check_inject_params(d.inject, f);
        } else {
            for (int c = 0; c < 3; ++c) PT_REQUIRE(d.x[c], "pt_serve_submit: null position array");
            if (d.inject_fixed_mu) {
                PT_REQUIRE(d.phase, "pt_serve_submit: fixed-mu injection needs gyrophases");
                PT_REQUIRE(d.mu0 >= -1.0 && d.mu0 <= 1.0, "injection_mu0 must be in [-1, 1]");
                PT_REQUIRE(std::isfinite(d.p_magnitude) && d.p_magnitude > 0,
                           "pt_serve_submit: p_magnitude must be positive");
            } else {
                for (int c = 0; c < 3; ++c) PT_REQUIRE(d.p[c], "pt_serve_submit: null momentum array");
            }
        }
        const size_t n = static_cast<size_t>(d.n);
        SReq r;
        r.id = srv->next_id++;
        r.d = d;
        r.nsave = d.out_kind != PT_OUT_NONE ? (d.total_steps - 1) / d.save_stride + 1 : 0;
        r.bytes = request_bytes(d, r.nsave);
        PT_REQUIRE(static_cast<int64_t>(r.bytes) <= srv->budget,
                   "pt_serve_submit: request needs " + std::to_string(r.bytes) + " bytes, more than the whole budget");
        if (!d.inject_on_device) {  // copy the host state (device injection draws it at admission)
            r.init.resize(d.inject_fixed_mu ? 4 * n : 6 * n);
            for (int c = 0; c < 3; ++c)
                for (size_t i = 0; i < n; ++i) {
                    const double v = d.x[c][i];
                    PT_REQUIRE(std::isfinite(v) && v >= f->g.lo[c] && v <= f->g.hi[c],
                               "pt_serve_submit: initial position outside the simulation box");
                    r.init[c * n + i] = v;
                }
            if (d.inject_fixed_mu) {
                std::memcpy(r.init.data() + 3 * n, d.phase, n * sizeof(double));
            } else {
                for (int c = 0; c < 3; ++c)
                    for (size_t i = 0; i < n; ++i) {
                        PT_REQUIRE(std::isfinite(d.p[c][i]), "pt_serve_submit: non-finite momentum");
                        r.init[(3 + c) * n + i] = d.p[c][i];
                    }
            }
        }
        for (int c = 0; c < 3; ++c) {
            r.d.x[c] = nullptr;
            r.d.p[c] = nullptr;
        }
        r.d.phase = nullptr;
        *request_id = r.id;
        srv->queue.push_back(r.id);
        srv->reqs.emplace(r.id, std::move(r));
    });
}

int pt_serve_step(pt_serve *srv, int64_t *n_running, int64_t *n_queued, int64_t *n_done) {
    return guarded([&] {
        PT_REQUIRE(srv, "pt_serve_step: null server");
        PT_HIP_CHECK(hipSetDevice(srv->ctx->device));
        // keep at most one tick in flight while this call enqueues the next
        while (srv->inflight.size() > 1) complete_oldest(srv);
        while (!srv->queue.empty() && static_cast<int32_t>(srv->active.size()) < srv->max_active) {
            SReq &r = srv->reqs.at(srv->queue.front());
            if (srv->reserved + static_cast<int64_t>(r.bytes) > srv->budget) break;  // FIFO: wait for memory
            if (!admit(srv, r)) break;  // no free block holds it yet: wait for releases
            srv->queue.pop_front();
            if (r.failed) {
                r.state = PT_REQ_DONE;
                srv->n_done += 1;
                srv->newly_done.push_back(r.id);
            } else {
                srv->active.push_back(r.id);
            }
        }
        const size_t before = srv->inflight.size();
        enqueue_tick(srv);
        // nothing new to run: finish what is in flight, so its requests are done
        if (srv->inflight.size() == before) complete_all(srv);
        if (n_running) *n_running = static_cast<int64_t>(srv->active.size());
        if (n_queued) *n_queued = static_cast<int64_t>(srv->queue.size());
        if (n_done) *n_done = srv->n_done;
    });
}

int pt_serve_take_done(pt_serve *srv, int64_t *ids, int64_t capacity, int64_t *count) {
    return guarded([&] {
        PT_REQUIRE(srv && count && (ids || capacity == 0) && capacity >= 0, "pt_serve_take_done: bad argument");
        const int64_t m = std::min<int64_t>(capacity, static_cast<int64_t>(srv->newly_done.size()));
        std::copy(srv->newly_done.begin(), srv->newly_done.begin() + m, ids);
        srv->newly_done.erase(srv->newly_done.begin(), srv->newly_done.begin() + m);
        *count = m;
    });
}

int pt_serve_drain(pt_serve *srv) {
    return guarded([&] {
        PT_REQUIRE(srv, "pt_serve_drain: null server");
        complete_all(srv);
    });
}

int pt_serve_poll(pt_serve *srv, int64_t request_id, int32_t *state) {
    return guarded([&] {
        PT_REQUIRE(srv && state, "pt_serve_poll: null argument");
        auto it = srv->reqs.find(request_id);
        PT_REQUIRE(it != srv->reqs.end(), "pt_serve_poll: unknown request id");
        *state = it->second.state;
    });
}

int pt_serve_fetch(pt_serve *srv, int64_t request_id, const void **cache_device, int64_t *nsave, int64_t *n,
                   int64_t *alive_final) {
    return guarded([&] {
        PT_REQUIRE(srv, "pt_serve_fetch: null server");
        auto it = srv->reqs.find(request_id);
        PT_REQUIRE(it != srv->reqs.end(), "pt_serve_fetch: unknown request id");
        SReq &r = it->second;
        if (r.state != PT_REQ_DONE) throw Error{PT_ERR_STATE, "pt_serve_fetch: request is not done"};
        if (r.failed) throw Error{PT_ERR_INVALID, "cannot use fixed-mu injection where local B is zero or non-finite"};
        if (cache_device) *cache_device = r.dev.out;
        if (nsave) *nsave = r.nsave;
        if (n) *n = r.d.n;
        if (alive_final) *alive_final = r.alive_final;
    });
}

int pt_serve_copy_cache(pt_serve *srv, int64_t request_id, void *host_dst, int64_t bytes) {
    return guarded([&] {
        PT_REQUIRE(srv && host_dst, "pt_serve_copy_cache: null argument");
        auto it = srv->reqs.find(request_id);
        PT_REQUIRE(it != srv->reqs.end(), "pt_serve_copy_cache: unknown request id");
        SReq &r = it->second;
        if (r.state != PT_REQ_DONE || r.failed || !r.dev.out)
            throw Error{PT_ERR_STATE, "pt_serve_copy_cache: no completed cache for this request"};
        const int64_t esz = r.d.out_kind == PT_OUT_PMAG ? 8 : 4;
        PT_REQUIRE(bytes == r.nsave * r.d.n * esz, "pt_serve_copy_cache: size mismatch");
        PT_HIP_CHECK(hipMemcpy(host_dst, r.dev.out, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
    });
}

int pt_serve_release(pt_serve *srv, int64_t request_id) {
    return guarded([&] {
        PT_REQUIRE(srv, "pt_serve_release: null server");
        auto it = srv->reqs.find(request_id);
        PT_REQUIRE(it != srv->reqs.end(), "pt_serve_release: unknown request id");
        SReq &r = it->second;
        if (r.state == PT_REQ_RUNNING) throw Error{PT_ERR_STATE, "pt_serve_release: request is still running"};
        if (r.state == PT_REQ_QUEUED) {
            srv->queue.erase(std::remove(srv->queue.begin(), srv->queue.end(), request_id), srv->queue.end());
        }
        if (r.mem) {
            // a DONE request's ticks have completed; its cache's other readers
            // finished before the caller releases it
            srv->arena.release(static_cast<size_t>(r.mem - srv->arena_base), r.bytes);
            srv->reserved -= static_cast<int64_t>(r.bytes);
            r.mem = nullptr;
        }
        srv->reqs.erase(it);
    });
}

int pt_serve_stats(pt_serve *srv, int64_t *bytes_reserved, int64_t *particle_steps, double *kernel_ms) {
    return guarded([&] {
        PT_REQUIRE(srv, "pt_serve_stats: null server");
        if (bytes_reserved) *bytes_reserved = srv->reserved;
        if (particle_steps) *particle_steps = srv->particle_steps;
        if (kernel_ms) *kernel_ms = srv->kernel_ms;
    });
}

}  // extern "C"
