// Cell tables: per-cell polynomial coefficients precomputed from a field's
// stored nodes (an acceleration structure the production pushers read
// instead of the nodes when it is present).
//
//  The production step loop keeps the current cell's 6 x 8 trilinear
//  coefficients in registers and rebuilds them only when the particle
//  changes cell.  From the stored nodes that rebuild is a wave-cooperative
//  pass (corner loads, decode, coefficient arithmetic, the coarse part of a
//  16-bit field) whose instruction cost does not depend on how many of the
//  wave's lanes need it, and at 10-100 TeV, where a particle crosses a cell
//  every few tens of pushes, a wave runs one such pass almost every push.  A
//  cell table holds the result of that arithmetic for every cell: a lane that
//  changed cell loads its 48 floats (192 bytes, one record) and multiplies
//  them by the request's physical factors.  The table is built with the same
//  device functions the pushers use (load_component_corners, cell_coeffs), so
//  the coefficients and every trajectory are bit-identical with and without
//  it.  Size: 192 bytes per cell (25.8 GB for a 512^3-cell field), against 16
//  or 32 bytes per node of the brick formats.
#include "pt_push_core.hpp"

namespace pt {
namespace {

using namespace core;

// One thread per (cell, component): the component's 8 field-unit
// coefficients of that cell, written to its slot of the cell's record.
template <int STORE>
__global__ void __launch_bounds__(256) k_build_cells(FieldView fv, int32_t nc0, int32_t nc1, int64_t n_items,
                                                     float *__restrict__ cells) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_items) return;
    const int comp = static_cast<int>(t % 6);
    const int64_t cell = t / 6;
    const int32_t i = static_cast<int32_t>(cell % nc0);
    const int32_t j = static_cast<int32_t>((cell / nc0) % nc1);
    const int32_t k = static_cast<int32_t>(cell / (static_cast<int64_t>(nc0) * nc1));
    float sv[8], c8[8];
    load_component_corners<STORE>(fv, i, j, k, comp, sv);
    cell_coeffs<STORE>(fv, i, j, k, comp, sv, fv.scale[comp], c8);
    float4 *dst = reinterpret_cast<float4 *>(cells + cell_record_index(fv.cb, i, j, k) * kCellFloats + comp * 8);
    dst[0] = make_float4(c8[0], c8[1], c8[2], c8[3]);
    dst[1] = make_float4(c8[4], c8[5], c8[6], c8[7]);
}

template <int STORE>
void launch_build(const pt_field *f, float *cells, hipStream_t st) {
    FieldView fv = make_field_view(f);
    for (int c = 0; c < 3; ++c) fv.cb[c] = f->cb[c];
    const int32_t nc0 = static_cast<int32_t>(f->g.n[0] - 1), nc1 = static_cast<int32_t>(f->g.n[1] - 1),
                  nc2 = static_cast<int32_t>(f->g.n[2] - 1);
    const int64_t items = static_cast<int64_t>(nc0) * nc1 * nc2 * 6;
    const int64_t blocks = (items + 255) / 256;
    k_build_cells<STORE><<<static_cast<unsigned>(blocks), 256, 0, st>>>(fv, nc0, nc1, items, cells);
    PT_HIP_CHECK(hipGetLastError());
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

}  // namespace
}  // namespace pt

using namespace pt;

int pt_field_cell_table(pt_field *field, int enable) {
    return guarded([&] {
        PT_REQUIRE(field, "pt_field_cell_table: null field");
        pt_ctx *ctx = field->ctx;
        if (!enable) {
            if (field->cells) {
                PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                dev_free(ctx, field->cells, field->cells_bytes);
                field->cells = nullptr;
                field->cells_bytes = 0;
            }
            return;
        }
        if (field->cells) return;
        PT_REQUIRE(field->storage == PT_STORE_F32_PLANAR || field->storage == PT_STORE_F32_BRICK ||
                       field->storage == PT_STORE_BF16_BRICK || field->storage == PT_STORE_F16_BRICK,
                   "pt_field_cell_table: needs an FP32 planar or brick field (the production pushers' storages)");
        int32_t cb[3];
        int64_t records = 512;
        for (int c = 0; c < 3; ++c) {
            const int64_t nc = field->g.n[c] - 1;
            PT_REQUIRE(nc >= 1, "pt_field_cell_table: a field axis has no cell");
            cb[c] = static_cast<int32_t>((nc + 7) / 8);
            records *= cb[c];
        }
        PT_REQUIRE(records <= INT32_MAX, "pt_field_cell_table: more than 2^31 cell records");
        const size_t bytes = static_cast<size_t>(records) * kCellFloats * sizeof(float);
        float *cells = static_cast<float *>(dev_alloc(ctx, bytes));
        for (int c = 0; c < 3; ++c) field->cb[c] = cb[c];
        try {
            switch (field->storage) {
                case PT_STORE_F32_PLANAR: launch_build<PT_STORE_F32_PLANAR>(field, cells, ctx->stream); break;
                case PT_STORE_F32_BRICK: launch_build<PT_STORE_F32_BRICK>(field, cells, ctx->stream); break;
                case PT_STORE_BF16_BRICK: launch_build<PT_STORE_BF16_BRICK>(field, cells, ctx->stream); break;
                case PT_STORE_F16_BRICK: launch_build<PT_STORE_F16_BRICK>(field, cells, ctx->stream); break;
                default: break;
            }
            PT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dev_free(ctx, cells, bytes);
            throw;
        }
        field->cells = cells;
        field->cells_bytes = bytes;
    });
}
