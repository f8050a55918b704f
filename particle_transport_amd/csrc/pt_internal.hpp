// Internal declarations shared by the libpt_hip.so translation units.
// Host handles, error plumbing, launch entry points of each kernel file and
// small device helpers (brick addressing, BF16/FP16 decode).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <string>

#include "../../include/pt_hip.h"

namespace pt {

// Physical constants of the reference (phase_space_gpu_runner.jl:8-10).
constexpr double kC = 2.99792458e8;
constexpr double kQe = 1.602176634e-19;
constexpr double kMp = 1.67262192369e-27;
constexpr double kGeVtoJ = 1e9 * 1.602176634e-19;

// ---------------------------------------------------------------- errors ----
void set_error(const std::string &msg);
struct Error {
    int code;
    std::string msg;
};
#define PT_HIP_CHECK(expr)                                                                 \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw ::pt::Error{PT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)
#define PT_REQUIRE(cond, msg)                                  \
    do {                                                       \
        if (!(cond)) throw ::pt::Error{PT_ERR_INVALID, (msg)}; \
    } while (0)

// ------------------------------------------------------------ geometry -----
// Grid of the reference: n nodes on [0, L] per axis, x_i = range(0, L, length=n)
// (phase_space_gpu_runner.jl:125-130), n-1 cells.  Bricks group nodes in 8^3 blocks.
struct Geom {
    int64_t n[3];      // nodes per axis
    double L;          // box length (m)
    double lo[3];      // xmin (0)
    double hi[3];      // xmax (L)
    double d[3];       // node spacing x[2]-x[1]
    int64_t nb[3];     // bricks per axis = ceil(n/8)
};

struct FieldScale {
    float scale[6];    // decoded value = (stored + coarse part, 16-bit bricks) * scale
};

// 16-bit bricks store each component as a coarse part plus a 16-bit residual.
// The coarse part is the trilinear interpolant of the field's own values at
// every 4th node per axis (the last coarse cell ends at the last node): per
// coarse cell and component, the FP32 coefficients of that interpolant as a
// polynomial a0 + a1 X + a2 Y + a3 Z + a4 XY + a5 XZ + a6 YZ + a7 XYZ in
// fine-node units X, Y, Z from the coarse cell's first node, in units of the
// component's scale_c (a power of two at least the largest residual, so the
// unit change is exact).  A node stores value / scale_c - coarse part in
// BF16/FP16, so the 16-bit rounding acts on the field's variation inside a
// 4^3-cell block instead of on its whole range.
constexpr int kCoarseShift = 2;   // coarse cell = 4 fine cells per axis
constexpr int kCoarseWidth = 1 << kCoarseShift;

// Coarse part at node (i, j, k) in FP64 from the stored coefficients: the
// node's coarse cell is floor(i/4) per axis, clamped to the last one (whose
// end node is the grid's last node).
__host__ __device__ inline double coarse_at_node(const float *__restrict__ table, const int32_t cn[3], int comp,
                                                 int64_t i, int64_t j, int64_t k) {
    int64_t I = i >> kCoarseShift, J = j >> kCoarseShift, K = k >> kCoarseShift;
    I = I < cn[0] - 1 ? I : cn[0] - 1;
    J = J < cn[1] - 1 ? J : cn[1] - 1;
    K = K < cn[2] - 1 ? K : cn[2] - 1;
    const double X = static_cast<double>(i - (I << kCoarseShift));
    const double Y = static_cast<double>(j - (J << kCoarseShift));
    const double Z = static_cast<double>(k - (K << kCoarseShift));
    const float *a = table + ((I + cn[0] * (J + cn[1] * K)) * 6 + comp) * 8;
    return static_cast<double>(a[0]) + static_cast<double>(a[1]) * X + static_cast<double>(a[2]) * Y +
           static_cast<double>(a[3]) * Z + static_cast<double>(a[4]) * X * Y + static_cast<double>(a[5]) * X * Z +
           static_cast<double>(a[6]) * Y * Z + static_cast<double>(a[7]) * X * Y * Z;
}

}  // namespace pt

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t bytes_in_use = 0;
    void *scratch = nullptr;     // reusable device scratch (grown on demand)
    size_t scratch_bytes = 0;
    void *pool = nullptr;        // cached per-call device buffers (pt_runtime.hip)
};

struct pt_field {
    pt_ctx *ctx = nullptr;
    pt::Geom g{};
    int storage = 0;
    void *data = nullptr;        // planar: 6 consecutive component arrays; brick: records
    size_t bytes = 0;
    pt::FieldScale sc{};         // brick storages only
    float *coarse = nullptr;     // 16-bit bricks: coarse-part coefficients [cell][6][8]
    size_t coarse_bytes = 0;
    int32_t cn[3] = {0, 0, 0};   // coarse cells per axis
    double B0 = 0.0;             // FP64 mean |B| of the source values
    // Optional cell table (pt_field_cell_table): per cell, the 6 x 8 FP32
    // polynomial coefficients the pushers would compute from the stored
    // nodes, in 8^3-cell bricks (cell_record_index); null when not built.
    float *cells = nullptr;
    size_t cells_bytes = 0;
    int32_t cb[3] = {0, 0, 0};   // cell bricks per axis
};

struct pt_ens {
    pt_ctx *ctx = nullptr;
    const pt_field *field = nullptr;
    int pusher = 0;
    int64_t n = 0;
    int periodic = 1;
    int64_t n_alive = 0;
    void *state = nullptr;       // see SplitState / RefState below
    size_t state_bytes = 0;
    uint8_t *alive = nullptr;
    int32_t *perm = nullptr;     // original index of the particle at each slot (after a sort)
    int64_t *deaths = nullptr;   // device per-step death counters
    int64_t deaths_cap = 0;
    // Morton re-sort: second copy of state/alive/perm (swapped with the live
    // one after each sort) and the radix-sort workspace, allocated at the
    // first sort and kept for the ensemble's lifetime.
    void *state_alt = nullptr;
    uint8_t *alive_alt = nullptr;
    int32_t *perm_alt = nullptr;
    void *sort_ws = nullptr;
    size_t sort_ws_bytes = 0;
    double kernel_ms = 0.0;
    int64_t launches = 0;
};

namespace pt {

// Split-precision particle state (structure of arrays, n elements each):
// double P2 = |p|^2 / (M_P c)^2, int32 cell[3], float off[3] in [0,1) cell
// units, float P[3] = p / (M_P c), float comp[3] = the rounding error the
// offset's last additions left (position = cell + off + comp, compensated
// summation).  56 bytes per particle.
struct SplitState {
    double *P2;
    int32_t *cell[3];
    float *off[3];
    float *P[3];
    float *comp[3];
};
constexpr size_t kSplitStateBytes = 56;
// FP64 reference-structured state: x in metres, p in kg m/s.
struct RefState {
    double *x[3];
    double *p[3];
};

inline SplitState split_state(pt_ens *e) {
    SplitState s;
    char *b = static_cast<char *>(e->state);
    size_t n = static_cast<size_t>(e->n);
    s.P2 = reinterpret_cast<double *>(b);
    b += n * 8;
    for (int c = 0; c < 3; ++c) s.cell[c] = reinterpret_cast<int32_t *>(b + c * n * 4);
    for (int c = 0; c < 3; ++c) s.off[c] = reinterpret_cast<float *>(b + (3 + c) * n * 4);
    for (int c = 0; c < 3; ++c) s.P[c] = reinterpret_cast<float *>(b + (6 + c) * n * 4);
    for (int c = 0; c < 3; ++c) s.comp[c] = reinterpret_cast<float *>(b + (9 + c) * n * 4);
    return s;
}
inline RefState ref_state(pt_ens *e) {
    RefState s;
    char *b = static_cast<char *>(e->state);
    size_t n = static_cast<size_t>(e->n);
    for (int c = 0; c < 3; ++c) s.x[c] = reinterpret_cast<double *>(b + c * n * 8);
    for (int c = 0; c < 3; ++c) s.p[c] = reinterpret_cast<double *>(b + (3 + c) * n * 8);
    return s;
}

// Per-call push parameters after host-side derivation.
struct PushArgs {
    int64_t first_step, n_steps, total_steps, stride;
    int64_t save_base;       // global 0-based index of the first save at/after first_step
    int out_kind;
    int out_f64;             // output element type: 1 double, 0 float
    void *out_a, *out_b;     // device output buffers [save][3][n] / [save][n]
    int64_t death_base;      // step whose deaths go to deaths[0] (the call's first step)
    int64_t out_cap;         // save slots in out buffers
    double dt, q, m;
    int steps_per_launch;
};

// Launchers (one per kernel file).  Each enqueues on e->ctx->stream and
// returns the number of kernel launches.
int64_t launch_push_ref_f64(pt_ens *e, const pt_field *f, const PushArgs &a);
int64_t launch_push_plain_f32(pt_ens *e, const pt_field *f, const PushArgs &a);
int64_t launch_push_fast(pt_ens *e, const pt_field *f, const PushArgs &a);

void launch_pack_field(pt_field *f, const void *const src[6], int src_f64, hipStream_t s);
void launch_mean_abs_b(pt_ctx *ctx, int64_t n, int src_f64, const void *bx, const void *by,
                       const void *bz, double *out_host);
void launch_split_from_si(pt_ens *e, const double *dx, const double *dp);
void morton_sort(pt_ens *e);  // pt_sort.hip
void launch_split_to_si(pt_ens *e, double *dx, double *dp, uint8_t *alive_out);
// pt_inject.hip: initial state of particles first..first+count-1 into out
// ([x1 x2 x3 p1 p2 p3][count], FP64 SI); *bad = 1 where fixed-mu meets zero |B|.
void launch_inject(const pt_field *f, const pt_inject_params &ip, int64_t first, int64_t count, double *out,
                   int *bad, hipStream_t s);
// The reference's checks on injection settings (pt_runtime.hip); throws.
void check_inject_params(const pt_inject_params &ip, const pt_field *field);

// Looks up (once per device, with `device` current) the device address of
// the fixed scratch array that pt_dmumu_accumulate_noalloc works in.
void noalloc_scratch_init(int device);
// Device scratch owned by the context, grown on demand (never shrinks).
void *ctx_scratch(pt_ctx *ctx, size_t bytes);
// Per-call device buffers cached by the context: pool_get returns a buffer of
// at least `bytes` (its size class in *granted), pool_put gives it back for a
// later call once the context's stream no longer uses it.  Saves the
// hipMalloc/hipFree pair, which synchronizes the device, on every call.
void *pool_get(pt_ctx *ctx, size_t bytes, size_t *granted);
void pool_put(pt_ctx *ctx, void *p, size_t granted);
// Tracked device allocations (counted in pt_ctx_bytes_in_use).
void *dev_alloc(pt_ctx *ctx, size_t bytes);
void dev_free(pt_ctx *ctx, void *p, size_t bytes);

// -------------------------------------------------------- device helpers ---
// Record of cell (i, j, k) in a cell table: cells grouped in 8^3 bricks
// (cb bricks per axis), 48 floats per record ([component][coefficient]).
__host__ __device__ inline int64_t cell_record_index(const int32_t cb[3], int64_t i, int64_t j, int64_t k) {
    const int64_t brick = ((k >> 3) * cb[1] + (j >> 3)) * cb[0] + (i >> 3);
    return (brick << 9) + (((k & 7) << 6) | ((j & 7) << 3) | (i & 7));
}
constexpr int kCellFloats = 48;
// Template tag of the pusher kernels that read a cell table instead of the
// stored nodes (not a storage kind of the C ABI).
constexpr int kStoreCells = 16;

__host__ __device__ inline int64_t brick_node_index(const int64_t nb[3], int64_t i, int64_t j,
                                                    int64_t k) {
    int64_t brick = ((k >> 3) * nb[1] + (j >> 3)) * nb[0] + (i >> 3);
    return (brick << 9) + (((k & 7) << 6) | ((j & 7) << 3) | (i & 7));
}

// Split coordinate of u = x / dx on an axis of nc cells: cell = floor(u), the
// FP32 offset u - cell (an offset that rounds up to 1 moves to the next cell
// at offset 0) and comp, the part of u - cell the offset's rounding dropped.
// A periodic axis holds the upper face x == L as the last cell at offset 1:
// the particle samples the field there as trilinear_sample clamps it
// (phase_space_gpu_runner.jl:247-276), and the first push that moves it
// outward takes it past offset 1, where the periodic wrap moves it to cell 0,
// as wrap_periodic_coordinate (:325-328) maps x > L to x - L.
__host__ __device__ inline void split_coord(double u, int32_t nc, bool periodic, int32_t &cell, float &off,
                                            float &comp) {
    const double fl = floor(u);
    const double frac = u - fl;
    off = static_cast<float>(frac);
    cell = static_cast<int32_t>(fl);
    comp = static_cast<float>(frac - static_cast<double>(off));
    if (off >= 1.0f) {
        off = 0.0f;
        cell += 1;
        comp = static_cast<float>(frac - 1.0);
    }
    if (periodic && cell == nc && off == 0.0f) {
        cell = nc - 1;
        off = 1.0f;
    }
}

__device__ inline float bf16_bits_to_float(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}
__device__ inline float f16_bits_to_float(uint16_t h) {
    return __half2float(__ushort_as_half(h));
}

}  // namespace pt
