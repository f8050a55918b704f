// libpt_h5.so: HDF5 caches and products in the reference's layouts, over the
// HDF5 C library.  Host-only code (no HIP): the cache writer thread of the
// runner calls pt_h5_write_box with page-locked chunks that the GPU has just
// filled, so the file writes overlap the pusher.
//
// Dimension order: the ABI takes Julia-order dims (first fastest); HDF5's
// dataspace lists the slowest dimension first, so every dims/start/count array
// is reversed on the way in and out.  This is the convention HDF5.jl uses
// (a Julia (n, 3, nsave) array is a (nsave, 3, n) dataspace), which the
// reference's own Python reader relies on (plot_energy_distribution.py:109-112).
#include "pt_h5.h"

#include <hdf5.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

struct pt_h5 {
    hid_t file = -1;
    int fd = -1;  // the sec2 driver's descriptor, for direct writes of cadence boxes
    std::mutex mu;
};

namespace {

thread_local std::string g_error;

struct Error {
    std::string msg;
};

// The innermost message on this thread's HDF5 error stack.
herr_t walk_cb(unsigned n, const H5E_error2_t *e, void *data) {
    auto *out = static_cast<std::string *>(data);
    if (n == 0 && e->desc) *out = e->desc;
    return 0;
}

[[noreturn]] void fail_h5(const std::string &what) {
    std::string detail;
    H5Ewalk2(H5E_DEFAULT, H5E_WALK_DOWNWARD, walk_cb, &detail);
    H5Eclear2(H5E_DEFAULT);
    throw Error{detail.empty() ? what : what + ": " + detail};
}

// HDF5 prints its error stack to stderr by default; every failure here is
// reported through the status and pt_h5_last_error instead.  The setting is
// per thread in a thread-safe HDF5 build.
void quiet() {
    thread_local bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        done = true;
    }
}

// Owner of one HDF5 identifier.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t), const std::string &what) : id(i), close(c) {
        if (id < 0) fail_h5(what);
    }
    Id(const Id &) = delete;
    Id &operator=(const Id &) = delete;
    Id(Id &&o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    Id &operator=(Id &&o) noexcept {
        if (this != &o) {
            if (id >= 0 && close) close(id);
            id = o.id;
            close = o.close;
            o.id = -1;
        }
        return *this;
    }
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

template <class F>
int guarded(F &&f) {
    quiet();
    try {
        f();
        return 0;
    } catch (const Error &e) {
        g_error = e.msg;
    } catch (const std::exception &e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown error";
    }
    return -1;
}

void need(bool ok, const std::string &msg) {
    if (!ok) throw Error{msg};
}

pt_h5 &handle(pt_h5 *f) {
    need(f && f->file >= 0, "invalid or closed file handle");
    return *f;
}

std::string path_of(const char *name) {
    need(name && *name, "empty dataset name");
    std::string s(name);
    while (!s.empty() && s.front() == '/') s.erase(s.begin());
    need(!s.empty(), "empty dataset name");
    return s;
}

hid_t mem_type(int dtype) {
    switch (dtype) {
        case PT_H5_F64: return H5T_NATIVE_DOUBLE;
        case PT_H5_F32: return H5T_NATIVE_FLOAT;
        case PT_H5_I64: return H5T_NATIVE_INT64;
        case PT_H5_I32: return H5T_NATIVE_INT32;
    }
    throw Error{"unknown element type " + std::to_string(dtype)};
}

hid_t file_type(int dtype) {
    switch (dtype) {
        case PT_H5_F64: return H5T_IEEE_F64LE;
        case PT_H5_F32: return H5T_IEEE_F32LE;
        case PT_H5_I64: return H5T_STD_I64LE;
        case PT_H5_I32: return H5T_STD_I32LE;
    }
    throw Error{"unknown element type " + std::to_string(dtype)};
}

std::vector<hsize_t> reversed(int rank, const int64_t *v, const char *what) {
    need(rank >= 0 && rank <= PT_H5_MAX_RANK, "rank must be 0.." + std::to_string(PT_H5_MAX_RANK));
    need(rank == 0 || v, std::string("missing ") + what);
    std::vector<hsize_t> h(rank);
    for (int k = 0; k < rank; ++k) {
        need(v[k] >= 0, std::string(what) + " must be non-negative");
        h[rank - 1 - k] = static_cast<hsize_t>(v[k]);
    }
    return h;
}

hid_t space_of(const std::vector<hsize_t> &h) {
    return h.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple(static_cast<int>(h.size()), h.data(), nullptr);
}

// Link creation with missing parent groups created, as HDF5.jl's
// file["a/b"] = x does.
hid_t lcpl() {
    const hid_t p = H5Pcreate(H5P_LINK_CREATE);
    if (p >= 0) H5Pset_create_intermediate_group(p, 1);
    return p;
}

bool exists(hid_t file, const std::string &path) {
    size_t pos = 0;
    while (true) {
        const size_t next = path.find('/', pos);
        const std::string prefix = path.substr(0, next);
        if (!prefix.empty() && prefix.back() != '/') {
            const htri_t r = H5Lexists(file, prefix.c_str(), H5P_DEFAULT);
            if (r < 0) fail_h5("cannot look up '" + prefix + "'");
            if (r == 0) return false;
        }
        if (next == std::string::npos) return true;
        pos = next + 1;
    }
}

void write_whole(hid_t file, const std::string &name, hid_t ftype, hid_t mtype, const std::vector<hsize_t> &h,
                 const void *src) {
    need(!exists(file, name), "dataset '" + name + "' already exists");
    Id space(space_of(h), H5Sclose, "cannot create dataspace");
    Id lc(lcpl(), H5Pclose, "cannot create property list");
    Id ds(H5Dcreate2(file, name.c_str(), ftype, space, lc, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
          "cannot create dataset '" + name + "'");
    hsize_t n = 1;
    for (hsize_t d : h) n *= d;
    if (n > 0 && H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, src) < 0)
        fail_h5("cannot write dataset '" + name + "'");
}

Id bool_type() {
    Id t(H5Tenum_create(H5T_NATIVE_INT8), H5Tclose, "cannot create Bool type");
    const int8_t f = 0, tr = 1;
    if (H5Tenum_insert(t, "FALSE", &f) < 0 || H5Tenum_insert(t, "TRUE", &tr) < 0) fail_h5("cannot create Bool type");
    return t;
}

struct Box {
    Id space;
    Id mem;
    hsize_t n = 0;
};

// File and memory dataspaces of a box of dataset `ds`.
void select_box(hid_t ds, const std::string &name, const int64_t *start, const int64_t *count, Box &b) {
    b.space = Id(H5Dget_space(ds), H5Sclose, "cannot read the dataspace of '" + name + "'");
    const int rank = H5Sget_simple_extent_ndims(b.space);
    need(rank >= 1, "dataset '" + name + "' is a scalar: boxes need rank >= 1");
    std::vector<hsize_t> dims(rank);
    H5Sget_simple_extent_dims(b.space, dims.data(), nullptr);
    const std::vector<hsize_t> st = reversed(rank, start, "start"), ct = reversed(rank, count, "count");
    b.n = 1;
    for (int k = 0; k < rank; ++k) {
        need(st[k] + ct[k] <= dims[k], "box outside dataset '" + name + "' (axis " + std::to_string(rank - k) +
                                           ": " + std::to_string(st[k]) + " + " + std::to_string(ct[k]) + " > " +
                                           std::to_string(dims[k]) + ")");
        b.n *= ct[k];
    }
    if (b.n == 0) return;
    if (H5Sselect_hyperslab(b.space, H5S_SELECT_SET, st.data(), nullptr, ct.data(), nullptr) < 0)
        fail_h5("cannot select a box of '" + name + "'");
    b.mem = Id(H5Screate_simple(rank, ct.data(), nullptr), H5Sclose, "cannot create dataspace");
}

// The sec2 driver's file descriptor, or -1 for another driver.
int vfd_of(hid_t file) {
    void *handle = nullptr;
    if (H5Fget_vfd_handle(file, H5P_DEFAULT, &handle) < 0 || !handle) {
        H5Eclear2(H5E_DEFAULT);
        return -1;
    }
    return *static_cast<int *>(handle);
}

// Threads for one direct write: PT_H5_WRITE_THREADS, else up to 8.
int write_threads() {
    static const int n = [] {
        const char *e = std::getenv("PT_H5_WRITE_THREADS");
        int v = e ? std::atoi(e) : 0;
        if (v <= 0) v = std::min(8, std::max(1, static_cast<int>(std::thread::hardware_concurrency())));
        return std::min(v, 64);
    }();
    return n;
}

void pwrite_all(int fd, const char *src, size_t bytes, off_t off, int &err) {
    while (bytes > 0) {
        const ssize_t w = ::pwrite(fd, src, std::min<size_t>(bytes, size_t{1} << 30), off);
        if (w < 0) {
            if (errno == EINTR) continue;
            err = errno;
            return;
        }
        src += w;
        off += w;
        bytes -= static_cast<size_t>(w);
    }
}

// A box that is one contiguous byte range of a contiguous, allocated dataset
// stored in the memory type is written with pwrite at its file offset, split
// over several threads: the page-cache copy of a cadence chunk (hundreds of
// MB) is the writer's cost, and one thread does not reach memory bandwidth.
// The box is contiguous when the axes below some axis j are whole, axis j is
// any range and the axes above j have count 1 (Julia order, first fastest).
// Returns false (the caller uses H5Dwrite) when any condition fails.
bool direct_write(pt_h5 &h, hid_t ds, hid_t mt, const int64_t *start, const int64_t *count, const void *src) {
    if (h.fd < 0) return false;
    Id dcpl(H5Dget_create_plist(ds), H5Pclose, "cannot read dataset storage");
    if (H5Pget_layout(dcpl) != H5D_CONTIGUOUS) return false;
    const haddr_t addr = H5Dget_offset(ds);
    if (addr == HADDR_UNDEF) {
        H5Eclear2(H5E_DEFAULT);
        return false;
    }
    Id ft(H5Dget_type(ds), H5Tclose, "cannot read dataset type");
    if (H5Tequal(ft, mt) <= 0) return false;
    Id space(H5Dget_space(ds), H5Sclose, "cannot read dataspace");
    const int rank = H5Sget_simple_extent_ndims(space);
    if (rank < 1) return false;
    std::vector<hsize_t> hd(rank);
    H5Sget_simple_extent_dims(space, hd.data(), nullptr);
    std::vector<int64_t> dims(rank);
    for (int k = 0; k < rank; ++k) dims[k] = static_cast<int64_t>(hd[rank - 1 - k]);
    int j = 0;
    while (j < rank - 1 && start[j] == 0 && count[j] == dims[j]) ++j;
    for (int k = j + 1; k < rank; ++k)
        if (count[k] != 1) return false;
    const size_t esz = H5Tget_size(ft);
    int64_t lin = 0, stride = 1, n = 1;
    for (int k = 0; k < rank; ++k) {
        lin += start[k] * stride;
        stride *= dims[k];
        n *= count[k];
    }
    const size_t bytes = static_cast<size_t>(n) * esz;
    const off_t off0 = static_cast<off_t>(addr) + static_cast<off_t>(lin) * static_cast<off_t>(esz);
    const size_t min_piece = size_t{16} << 20;
    const int pieces = static_cast<int>(std::max<size_t>(1, std::min<size_t>(write_threads(), bytes / min_piece)));
    const char *base = static_cast<const char *>(src);
    std::vector<int> errs(pieces, 0);
    if (pieces == 1) {
        pwrite_all(h.fd, base, bytes, off0, errs[0]);
    } else {
        const size_t step = (bytes + pieces - 1) / pieces;
        std::vector<std::thread> ts;
        for (int t = 0; t < pieces; ++t) {
            const size_t a = std::min(bytes, t * step), e = std::min(bytes, a + step);
            ts.emplace_back(pwrite_all, h.fd, base + a, e - a, off0 + static_cast<off_t>(a), std::ref(errs[t]));
        }
        for (auto &t : ts) t.join();
    }
    for (int e : errs)
        if (e) throw Error{std::string("cannot write the file: ") + std::strerror(e)};
    return true;
}

}  // namespace

extern "C" {

const char *pt_h5_last_error(void) { return g_error.c_str(); }

int pt_h5_create(const char *path, pt_h5 **out) {
    return guarded([&] {
        need(path && *path && out, "missing path or output handle");
        *out = nullptr;
        Id fapl(H5Pcreate(H5P_FILE_ACCESS), H5Pclose, "cannot create property list");
        H5Pset_fclose_degree(fapl, H5F_CLOSE_STRONG);
        // No raw-data sieve buffer: cadence boxes bypass the library (direct
        // writes below), so the library must hold no raw bytes of its own.
        H5Pset_sieve_buf_size(fapl, 0);
        const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
        if (f < 0) fail_h5(std::string("cannot create '") + path + "'");
        auto *h = new pt_h5;
        h->file = f;
        h->fd = vfd_of(f);
        *out = h;
    });
}

int pt_h5_open(const char *path, int writable, pt_h5 **out) {
    return guarded([&] {
        need(path && *path && out, "missing path or output handle");
        *out = nullptr;
        Id fapl(H5Pcreate(H5P_FILE_ACCESS), H5Pclose, "cannot create property list");
        H5Pset_fclose_degree(fapl, H5F_CLOSE_STRONG);
        if (writable) H5Pset_sieve_buf_size(fapl, 0);
        const hid_t f = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, fapl);
        if (f < 0) fail_h5(std::string("cannot open '") + path + "'");
        auto *h = new pt_h5;
        h->file = f;
        h->fd = writable ? vfd_of(f) : -1;
        *out = h;
    });
}

int pt_h5_close(pt_h5 *f) {
    return guarded([&] {
        if (!f) return;
        hid_t id;
        {
            std::lock_guard<std::mutex> lk(f->mu);
            id = f->file;
            f->file = -1;
        }
        delete f;
        if (id >= 0 && H5Fclose(id) < 0) fail_h5("cannot close file");
    });
}

int pt_h5_flush(pt_h5 *f) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        std::lock_guard<std::mutex> lk(h.mu);
        if (H5Fflush(h.file, H5F_SCOPE_GLOBAL) < 0) fail_h5("cannot flush file");
    });
}

int pt_h5_exists(pt_h5 *f, const char *name, int *out) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        need(out, "missing output");
        std::lock_guard<std::mutex> lk(h.mu);
        *out = exists(h.file, path_of(name)) ? 1 : 0;
    });
}

int pt_h5_create_group(pt_h5 *f, const char *name) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        std::lock_guard<std::mutex> lk(h.mu);
        need(!exists(h.file, p), "'" + p + "' already exists");
        Id lc(lcpl(), H5Pclose, "cannot create property list");
        Id g(H5Gcreate2(h.file, p.c_str(), lc, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "cannot create group '" + p + "'");
    });
}

int pt_h5_create_dataset(pt_h5 *f, const char *name, int dtype, int rank, const int64_t *dims) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        const hid_t ft = file_type(dtype);
        need(rank >= 1, "create_dataset needs rank >= 1");
        const std::vector<hsize_t> hd = reversed(rank, dims, "dims");
        std::lock_guard<std::mutex> lk(h.mu);
        need(!exists(h.file, p), "dataset '" + p + "' already exists");
        Id space(space_of(hd), H5Sclose, "cannot create dataspace");
        Id lc(lcpl(), H5Pclose, "cannot create property list");
        // Contiguous storage allocated now (its file offset is then known, so
        // cadence boxes can be written directly) and never pre-filled: every
        // element is written by the cadence writer.
        Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "cannot create property list");
        if (H5Pset_layout(dcpl, H5D_CONTIGUOUS) < 0 || H5Pset_alloc_time(dcpl, H5D_ALLOC_TIME_EARLY) < 0 ||
            H5Pset_fill_time(dcpl, H5D_FILL_TIME_NEVER) < 0)
            fail_h5("cannot set dataset storage");
        Id ds(H5Dcreate2(h.file, p.c_str(), ft, space, lc, dcpl, H5P_DEFAULT), H5Dclose,
              "cannot create dataset '" + p + "'");
    });
}

int pt_h5_write_box(pt_h5 *f, const char *name, int dtype, const int64_t *start, const int64_t *count,
                    const void *src) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        const hid_t mt = mem_type(dtype);
        std::lock_guard<std::mutex> lk(h.mu);
        Id ds(H5Dopen2(h.file, p.c_str(), H5P_DEFAULT), H5Dclose, "no dataset '" + p + "'");
        Box b;
        select_box(ds, p, start, count, b);
        if (b.n == 0) return;
        need(src, "missing source buffer");
        if (direct_write(h, ds, mt, start, count, src)) return;
        if (H5Dwrite(ds, mt, b.mem, b.space, H5P_DEFAULT, src) < 0) fail_h5("cannot write a box of '" + p + "'");
    });
}

int pt_h5_read_box(pt_h5 *f, const char *name, int dtype, const int64_t *start, const int64_t *count, void *dst) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        const hid_t mt = mem_type(dtype);
        std::lock_guard<std::mutex> lk(h.mu);
        Id ds(H5Dopen2(h.file, p.c_str(), H5P_DEFAULT), H5Dclose, "no dataset '" + p + "'");
        Box b;
        select_box(ds, p, start, count, b);
        if (b.n == 0) return;
        need(dst, "missing destination buffer");
        if (H5Dread(ds, mt, b.mem, b.space, H5P_DEFAULT, dst) < 0) fail_h5("cannot read a box of '" + p + "'");
    });
}

int pt_h5_write_array(pt_h5 *f, const char *name, int dtype, int rank, const int64_t *dims, const void *src) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        const hid_t ft = file_type(dtype), mt = mem_type(dtype);
        const std::vector<hsize_t> hd = reversed(rank, dims, "dims");
        need(src || std::any_of(hd.begin(), hd.end(), [](hsize_t d) { return d == 0; }), "missing source buffer");
        std::lock_guard<std::mutex> lk(h.mu);
        write_whole(h.file, p, ft, mt, hd, src);
    });
}

int pt_h5_write_string(pt_h5 *f, const char *name, const char *utf8) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        need(utf8, "missing string");
        const std::string s(utf8);
        std::lock_guard<std::mutex> lk(h.mu);
        Id t(H5Tcopy(H5T_C_S1), H5Tclose, "cannot create string type");
        if (H5Tset_size(t, std::max<size_t>(s.size(), 1)) < 0 || H5Tset_cset(t, H5T_CSET_UTF8) < 0 ||
            H5Tset_strpad(t, H5T_STR_NULLTERM) < 0)
            fail_h5("cannot create string type");
        std::string buf = s.empty() ? std::string(1, '\0') : s;
        write_whole(h.file, p, t, t, {}, buf.data());
    });
}

int pt_h5_write_strings(pt_h5 *f, const char *name, int64_t count, const char *const *utf8) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        need(count >= 0 && (count == 0 || utf8), "missing strings");
        for (int64_t k = 0; k < count; ++k) need(utf8[k], "missing string " + std::to_string(k));
        std::lock_guard<std::mutex> lk(h.mu);
        Id t(H5Tcopy(H5T_C_S1), H5Tclose, "cannot create string type");
        if (H5Tset_size(t, H5T_VARIABLE) < 0 || H5Tset_cset(t, H5T_CSET_UTF8) < 0) fail_h5("cannot create string type");
        write_whole(h.file, p, t, t, {static_cast<hsize_t>(count)}, utf8);
    });
}

int pt_h5_write_bool(pt_h5 *f, const char *name, int value) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        std::lock_guard<std::mutex> lk(h.mu);
        Id t = bool_type();
        const int8_t v = value ? 1 : 0;
        write_whole(h.file, p, t, t, {}, &v);
    });
}

int pt_h5_info(pt_h5 *f, const char *name, int *kind, int *dtype, int *rank, int64_t *dims) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        need(kind && dtype && rank && dims, "missing outputs");
        std::lock_guard<std::mutex> lk(h.mu);
        need(exists(h.file, p), "no dataset '" + p + "'");
        Id ds(H5Dopen2(h.file, p.c_str(), H5P_DEFAULT), H5Dclose, "'" + p + "' is not a dataset");
        Id t(H5Dget_type(ds), H5Tclose, "cannot read the type of '" + p + "'");
        Id space(H5Dget_space(ds), H5Sclose, "cannot read the dataspace of '" + p + "'");
        const int r = H5Sget_simple_extent_ndims(space);
        need(r >= 0 && r <= PT_H5_MAX_RANK, "dataset '" + p + "' has an unsupported rank");
        std::vector<hsize_t> hd(r);
        if (r > 0) H5Sget_simple_extent_dims(space, hd.data(), nullptr);
        *rank = r;
        for (int k = 0; k < r; ++k) dims[k] = static_cast<int64_t>(hd[r - 1 - k]);
        const H5T_class_t cls = H5Tget_class(t);
        const size_t sz = H5Tget_size(t);
        *dtype = -1;
        if (cls == H5T_STRING) {
            *kind = PT_H5_STRING;
        } else if (cls == H5T_ENUM) {
            *kind = PT_H5_BOOL;
        } else if (cls == H5T_INTEGER) {
            *kind = PT_H5_NUMERIC;
            *dtype = sz > 4 ? PT_H5_I64 : PT_H5_I32;
        } else if (cls == H5T_FLOAT) {
            *kind = PT_H5_NUMERIC;
            *dtype = sz > 4 ? PT_H5_F64 : PT_H5_F32;
        } else {
            throw Error{"dataset '" + p + "' has an unsupported type"};
        }
    });
}

int pt_h5_read_array(pt_h5 *f, const char *name, int dtype, void *dst, int64_t capacity) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        const hid_t mt = mem_type(dtype);
        std::lock_guard<std::mutex> lk(h.mu);
        Id ds(H5Dopen2(h.file, p.c_str(), H5P_DEFAULT), H5Dclose, "no dataset '" + p + "'");
        Id space(H5Dget_space(ds), H5Sclose, "cannot read the dataspace of '" + p + "'");
        const hssize_t n = H5Sget_simple_extent_npoints(space);
        need(n >= 0 && n <= capacity, "dataset '" + p + "' holds " + std::to_string(n) + " elements, more than " +
                                          std::to_string(capacity));
        if (n == 0) return;
        need(dst, "missing destination buffer");
        Id t(H5Dget_type(ds), H5Tclose, "cannot read the type of '" + p + "'");
        if (H5Tget_class(t) == H5T_ENUM) {
            // Bool: read the enum's integer values, then widen.
            Id nt(H5Tget_native_type(t, H5T_DIR_ASCEND), H5Tclose, "cannot read the type of '" + p + "'");
            const size_t sz = H5Tget_size(nt);
            need(sz <= 8, "enum dataset '" + p + "' has an unsupported size");
            std::vector<unsigned char> raw(static_cast<size_t>(n) * sz);
            if (H5Dread(ds, nt, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw.data()) < 0) fail_h5("cannot read '" + p + "'");
            for (hssize_t k = 0; k < n; ++k) {
                int64_t v = 0;
                bool nonzero = false;
                for (size_t b = 0; b < sz; ++b) nonzero |= raw[k * sz + b] != 0;
                v = nonzero ? 1 : 0;
                switch (dtype) {
                    case PT_H5_F64: static_cast<double *>(dst)[k] = static_cast<double>(v); break;
                    case PT_H5_F32: static_cast<float *>(dst)[k] = static_cast<float>(v); break;
                    case PT_H5_I64: static_cast<int64_t *>(dst)[k] = v; break;
                    case PT_H5_I32: static_cast<int32_t *>(dst)[k] = static_cast<int32_t>(v); break;
                }
            }
            return;
        }
        need(H5Tget_class(t) == H5T_INTEGER || H5Tget_class(t) == H5T_FLOAT, "dataset '" + p + "' is not numeric");
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, dst) < 0) fail_h5("cannot read '" + p + "'");
    });
}

int pt_h5_read_string(pt_h5 *f, const char *name, int64_t index, char *buf, int64_t cap, int64_t *len) {
    return guarded([&] {
        pt_h5 &h = handle(f);
        const std::string p = path_of(name);
        need(buf && cap > 0 && len, "missing output buffer");
        std::lock_guard<std::mutex> lk(h.mu);
        Id ds(H5Dopen2(h.file, p.c_str(), H5P_DEFAULT), H5Dclose, "no dataset '" + p + "'");
        Id t(H5Dget_type(ds), H5Tclose, "cannot read the type of '" + p + "'");
        need(H5Tget_class(t) == H5T_STRING, "dataset '" + p + "' is not a string");
        Id space(H5Dget_space(ds), H5Sclose, "cannot read the dataspace of '" + p + "'");
        const hssize_t n = H5Sget_simple_extent_npoints(space);
        need(index >= 0 && index < n, "string index " + std::to_string(index) + " outside '" + p + "'");
        std::string s;
        if (H5Tis_variable_str(t) > 0) {
            Id mt(H5Tcopy(H5T_C_S1), H5Tclose, "cannot create string type");
            H5Tset_size(mt, H5T_VARIABLE);
            H5Tset_cset(mt, H5Tget_cset(t));
            std::vector<char *> ptrs(static_cast<size_t>(n), nullptr);
            if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptrs.data()) < 0) fail_h5("cannot read '" + p + "'");
            if (ptrs[index]) s = ptrs[index];
            H5Dvlen_reclaim(mt, space, H5P_DEFAULT, ptrs.data());
        } else {
            const size_t sz = H5Tget_size(t);
            std::vector<char> raw(static_cast<size_t>(n) * sz);
            if (H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw.data()) < 0) fail_h5("cannot read '" + p + "'");
            const char *e = raw.data() + static_cast<size_t>(index) * sz;
            size_t l = 0;
            while (l < sz && e[l] != '\0') ++l;
            s.assign(e, l);
            if (H5Tget_strpad(t) == H5T_STR_SPACEPAD)
                while (!s.empty() && s.back() == ' ') s.pop_back();
        }
        *len = static_cast<int64_t>(s.size());
        const size_t m = std::min<size_t>(s.size(), static_cast<size_t>(cap - 1));
        std::memcpy(buf, s.data(), m);
        buf[m] = '\0';
    });
}

}  // extern "C"
