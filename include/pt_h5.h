/* pt_h5.h - C ABI of libpt_h5.so: HDF5 files in the reference's layouts.
 *
 * The reference writes its caches and products with HDF5.jl: a Julia array of
 * size (a, b, c) is stored column-major, so C and h5py see dimensions
 * (c, b, a).  Every dims/start/count array in this ABI is in Julia order (first
 * entry fastest in memory), as a Julia caller passes size(A); memory buffers
 * are column-major in that order (equivalently C-order in the reversed order,
 * which is how numpy sees them).
 *
 * Scalars are written the two ways the reference writes them: a length-1
 * vector (`Float64[x]`, `Int[x]`, rank 1, dims {1}) or a true scalar
 * (`Int(x)`, rank 0).  Strings are scalar fixed-length UTF-8 strings of the
 * string's byte length; string vectors are variable-length UTF-8; Bool is an
 * 8-bit enum {FALSE = 0, TRUE = 1}.  The exact string and Bool datatypes
 * HDF5.jl emits are not pinned by any fixture; the reference's readers accept
 * Bool, integers and strings for flags (h5_scalar_bool_or,
 * multimode_multienergy_cache_pipeline.jl:1072-1080) and either scalar form
 * (h5_scalar_value_or, :1041-1049).
 *
 * Every function returns 0 on success or a negative status; the message of
 * the calling thread's last failure is pt_h5_last_error().  A handle may be
 * used from any thread, one call at a time per handle.
 */
#ifndef PT_H5_H
#define PT_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_h5 pt_h5;

/* Element types.  PT_H5_F64 and PT_H5_F32 have the values of PT_F64 and
 * PT_F32 in pt_hip.h. */
enum pt_h5_dtype { PT_H5_F64 = 0, PT_H5_F32 = 1, PT_H5_I64 = 2, PT_H5_I32 = 3 };

/* What a dataset holds (pt_h5_info). */
enum pt_h5_kind { PT_H5_NUMERIC = 0, PT_H5_STRING = 1, PT_H5_BOOL = 2 };

#define PT_H5_MAX_RANK 8

const char *pt_h5_last_error(void);

/* h5open(path, "w") (truncates) / h5open(path, "r") or "r+". */
int pt_h5_create(const char *path, pt_h5 **out);
int pt_h5_open(const char *path, int writable, pt_h5 **out);
int pt_h5_close(pt_h5 *f);
int pt_h5_flush(pt_h5 *f);

/* haskey(file, name) for a '/'-separated path of groups and a final link. */
int pt_h5_exists(pt_h5 *f, const char *name, int *out);
/* create_group(file, name); missing parent groups are created. */
int pt_h5_create_group(pt_h5 *f, const char *name);

/* create_dataset(file, name, T, dims): contiguous storage, written later in
 * boxes (create_phase_space_writer, phase_space_gpu_runner.jl:406-411;
 * create_mu_cache_writer, multimode_multienergy_cache_pipeline.jl:1312-1316). */
int pt_h5_create_dataset(pt_h5 *f, const char *name, int dtype, int rank, const int64_t *dims);

/* dataset[start+1 : start+count, ...] = src, and the converse read.  start is
 * 0-based; src/dst hold prod(count) elements of `dtype`, column-major; the
 * library converts between `dtype` and the dataset's type.  A box that spans
 * whole leading axes and a range of the last axis is one contiguous file
 * region (write_phase_space_step!, phase_space_gpu_runner.jl:413-421, writes
 * one save index at a time; here any run of save indices at once). */
int pt_h5_write_box(pt_h5 *f, const char *name, int dtype, const int64_t *start, const int64_t *count,
                    const void *src);
int pt_h5_read_box(pt_h5 *f, const char *name, int dtype, const int64_t *start, const int64_t *count, void *dst);

/* file[name] = A: create and write a whole numeric dataset (rank 0: scalar). */
int pt_h5_write_array(pt_h5 *f, const char *name, int dtype, int rank, const int64_t *dims, const void *src);
/* file[name] = "text" (scalar string), = ["a", "b"] (string vector), = true. */
int pt_h5_write_string(pt_h5 *f, const char *name, const char *utf8);
int pt_h5_write_strings(pt_h5 *f, const char *name, int64_t count, const char *const *utf8);
int pt_h5_write_bool(pt_h5 *f, const char *name, int value);

/* Kind, stored element type (numeric datasets; strings and Bool report -1),
 * rank and Julia-order dims (up to PT_H5_MAX_RANK) of a dataset. */
int pt_h5_info(pt_h5 *f, const char *name, int *kind, int *dtype, int *rank, int64_t *dims);
/* read(file[name]) of a numeric or Bool dataset into dst (capacity in
 * elements, converted to dtype). */
int pt_h5_read_array(pt_h5 *f, const char *name, int dtype, void *dst, int64_t capacity);
/* Element `index` of a string dataset (0 for a scalar string) into buf
 * (NUL-terminated, truncated to cap - 1 bytes); *len gets the full length. */
int pt_h5_read_string(pt_h5 *f, const char *name, int64_t index, char *buf, int64_t cap, int64_t *len);

#ifdef __cplusplus
}
#endif

#endif /* PT_H5_H */
