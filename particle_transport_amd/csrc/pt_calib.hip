// Vector-ALU issue-rate calibration for bench.py (libpt_calib.so).
//
// Not part of the drop-in boundary (include/pt_hip.h): a separate library of
// small kernels that measure how many wave64 instructions per second the
// MI355X issues for the instruction forms the pusher's hot loop uses, so the
// bench can state its VALU roof from a measurement on the same box beside
// the published FP32 peak.  Each kernel runs eight independent dependency
// chains per lane (enough work per wave that issue, not latency, bounds it)
// over `iters` iterations of 8 x 8 instructions of one form:
//
//   PT_CALIB_FMA_VGPR   v_fma_f32 with every operand in a VGPR
//   PT_CALIB_FMA_SGPR   v_fma_f32 with one operand in an SGPR
//   PT_CALIB_ADD_VGPR   v_add_f32 with both operands in VGPRs
//   PT_CALIB_RSQ        v_rsq_f32 (a transcendental)
//   PT_CALIB_FMA_F64    v_fma_f64 (4 chains, 8 x 4 instructions per iteration)
//
// Every lane writes one float to out[global lane], so `out` holds
// blocks x 256 floats.  Built like the library (packed FP32 disabled), so
// no instruction of these kernels is a packed one.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

enum { PT_CALIB_FMA_VGPR = 0, PT_CALIB_FMA_SGPR = 1, PT_CALIB_ADD_VGPR = 2, PT_CALIB_RSQ = 3, PT_CALIB_FMA_F64 = 4 };

template <int KIND>
__global__ void __launch_bounds__(256) k_calib(float *__restrict__ out, int iters, float a, float b) {
    // per-lane values the compiler cannot fold into constants (VGPR operands)
    const float va = a + static_cast<float>(threadIdx.x) * 1e-9f;
    const float vb = b + static_cast<float>(threadIdx.x) * 1e-9f;
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = static_cast<float>(threadIdx.x) * 1e-3f + static_cast<float>(i + 1);
    double dacc[4] = {1.0, 2.0, 3.0, 4.0};
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if constexpr (KIND == PT_CALIB_FMA_F64) {
#pragma unroll
                for (int i = 0; i < 4; ++i) dacc[i] = fma(dacc[i], static_cast<double>(va), static_cast<double>(vb));
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if constexpr (KIND == PT_CALIB_FMA_VGPR) acc[i] = fmaf(acc[i], va, vb);
                    if constexpr (KIND == PT_CALIB_FMA_SGPR) acc[i] = fmaf(acc[i], a, vb);
                    if constexpr (KIND == PT_CALIB_ADD_VGPR) acc[i] = acc[i] + va;
                    if constexpr (KIND == PT_CALIB_RSQ) acc[i] = __builtin_amdgcn_rsqf(acc[i]);
                }
            }
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += acc[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) s += static_cast<float>(dacc[i]);
    out[static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

template <int KIND>
void launch(float *out, int blocks, int iters, hipStream_t st) {
    k_calib<KIND><<<blocks, 256, 0, st>>>(out, iters, 0.999f, 1e-3f);
}

}  // namespace

// Runs kernel `kind` twice (a 4-iteration warmup, then `iters` iterations) on
// `blocks` workgroups of 256 lanes on the current device and stores the
// second launch's duration (HIP events on a stream of its own) in *ms.
// Wave64 instructions issued by the timed launch: blocks x 4 x iters x 64
// (x 32 for PT_CALIB_FMA_F64).  Returns 0, or a hipError_t value.
extern "C" int pt_calib_valu(int kind, int blocks, int iters, float *ms) {
    if (kind < 0 || kind > PT_CALIB_FMA_F64 || blocks <= 0 || iters <= 0 || !ms) return static_cast<int>(hipErrorInvalidValue);
    float *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(&out, sizeof(float) * static_cast<size_t>(blocks) * 256);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
        for (int rep = 0; rep < 2 && err == hipSuccess; ++rep) {
            const int n = rep == 0 ? 4 : iters;
            if (rep == 1) err = hipEventRecord(e0, st);
            switch (kind) {
                case PT_CALIB_FMA_VGPR: launch<PT_CALIB_FMA_VGPR>(out, blocks, n, st); break;
                case PT_CALIB_FMA_SGPR: launch<PT_CALIB_FMA_SGPR>(out, blocks, n, st); break;
                case PT_CALIB_ADD_VGPR: launch<PT_CALIB_ADD_VGPR>(out, blocks, n, st); break;
                case PT_CALIB_RSQ: launch<PT_CALIB_RSQ>(out, blocks, n, st); break;
                default: launch<PT_CALIB_FMA_F64>(out, blocks, n, st); break;
            }
            if (err == hipSuccess) err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, st);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return static_cast<int>(err);
}
