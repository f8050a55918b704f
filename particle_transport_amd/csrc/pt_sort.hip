// Morton re-sort of a split-precision ensemble.
//
// Key = Morton interleave of the particle's brick coordinates (cell >> 3,
// 10 bits per axis); dead particles sort last.  hipCUB's radix sort orders
// (key, slot) pairs; a gather then rebuilds the state arrays in the new
// order and composes the permutation that maps each slot back to the
// particle's original index, which the pushers use as the output column.
#include <hipcub/hipcub.hpp>

#include <utility>

#include "pt_internal.hpp"

namespace pt {
namespace {

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton_keys(SplitState s, const uint8_t *__restrict__ alive, int64_t n,
                              uint32_t *__restrict__ keys, int32_t *__restrict__ slots) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t key = 0xFFFFFFFFu;
    if (alive[i]) {
        uint32_t b[3];
        for (int c = 0; c < 3; ++c) {
            const int32_t cell = s.cell[c][i];
            b[c] = static_cast<uint32_t>(cell < 0 ? 0 : cell) >> 3;
        }
        key = spread10(b[0]) | (spread10(b[1]) << 1) | (spread10(b[2]) << 2);
    }
    keys[i] = key;
    slots[i] = static_cast<int32_t>(i);
}

__global__ void k_gather_state(SplitState src, SplitState dst, const uint8_t *__restrict__ alive_src,
                               uint8_t *__restrict__ alive_dst, const int32_t *__restrict__ perm_src,
                               int32_t *__restrict__ perm_dst, const int32_t *__restrict__ order, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = order[i];
    dst.P2[i] = src.P2[j];
    for (int c = 0; c < 3; ++c) {
        dst.cell[c][i] = src.cell[c][j];
        dst.off[c][i] = src.off[c][j];
        dst.comp[c][i] = src.comp[c][j];
        dst.P[c][i] = src.P[c][j];
    }
    alive_dst[i] = alive_src[j];
    perm_dst[i] = perm_src ? perm_src[j] : j;
}

}  // namespace

void morton_sort(pt_ens *e) {
    PT_REQUIRE(e->pusher != PT_PUSH_REF_F64, "pt_ens_sort: the FP64 reference pusher keeps injection order");
    pt_ctx *ctx = e->ctx;
    hipStream_t st = ctx->stream;
    const int64_t n = e->n;
    const size_t N = static_cast<size_t>(n);
    // Workspace owned by the ensemble (never the context scratch, which holds
    // host-bound output staging while a segmented advance is in progress):
    // keys/slots in and out, then the radix-sort temporary storage.
    size_t temp_bytes = 0;
    PT_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, static_cast<uint32_t *>(nullptr),
                                                    static_cast<uint32_t *>(nullptr), static_cast<int32_t *>(nullptr),
                                                    static_cast<int32_t *>(nullptr), static_cast<int>(n), 0, 32, st));
    const size_t head = ((16 * N + 255) / 256) * 256;
    if (e->sort_ws_bytes < head + temp_bytes) {
        PT_HIP_CHECK(hipStreamSynchronize(st));
        dev_free(ctx, e->sort_ws, e->sort_ws_bytes);
        e->sort_ws = nullptr;
        e->sort_ws_bytes = 0;
        e->sort_ws = dev_alloc(ctx, head + temp_bytes);
        e->sort_ws_bytes = head + temp_bytes;
    }
    if (!e->state_alt) {
        e->state_alt = dev_alloc(ctx, e->state_bytes);
        e->alive_alt = static_cast<uint8_t *>(dev_alloc(ctx, N));
        e->perm_alt = static_cast<int32_t *>(dev_alloc(ctx, N * 4));
    }
    char *scr = static_cast<char *>(e->sort_ws);
    auto *keys_in = reinterpret_cast<uint32_t *>(scr);
    auto *keys_out = keys_in + N;
    auto *slots_in = reinterpret_cast<int32_t *>(keys_out + N);
    auto *slots_out = slots_in + N;
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    SplitState src = split_state(e);
    k_morton_keys<<<blocks, 256, 0, st>>>(src, e->alive, n, keys_in, slots_in);
    PT_HIP_CHECK(hipGetLastError());
    PT_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(scr + head, temp_bytes, keys_in, keys_out, slots_in, slots_out,
                                                    static_cast<int>(n), 0, 32, st));
    // Gather into the spare copy, then swap it in.  Stream order makes the
    // old arrays free for the next sort without a host synchronisation.
    std::swap(e->state, e->state_alt);
    SplitState dst = split_state(e);
    k_gather_state<<<blocks, 256, 0, st>>>(src, dst, e->alive, e->alive_alt, e->perm, e->perm_alt, slots_out, n);
    PT_HIP_CHECK(hipGetLastError());
    std::swap(e->alive, e->alive_alt);
    if (!e->perm) {
        e->perm = e->perm_alt;
        e->perm_alt = static_cast<int32_t *>(dev_alloc(ctx, N * 4));
    } else {
        std::swap(e->perm, e->perm_alt);
    }
}

}  // namespace pt
