// OPS-layout HDF5 output and restart (include/osb_io.h).  Host code only,
// built with g++ against the HDF5 C library; the HIP library does not
// depend on it.
#include "../../include/osb.h"
#include "../../include/osb_io.h"

#include <hdf5.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

std::mutex g_h5_mutex;              // the HDF5 library is built without thread safety
thread_local std::string g_io_error;

int fail(int code, const std::string& m) {
    g_io_error = m;
    return code;
}

// closes an HDF5 handle with the matching routine at scope exit
struct H {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H() = default;
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    H(const H&) = delete;
    H& operator=(const H&) = delete;
    ~H() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

// int32 vector attribute (H5T_STD_I32LE, like h5py's dtype="int32")
bool attr_ints(hid_t obj, const char* name, const int32_t* v, hsize_t n) {
    H sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return false;
    H at(H5Acreate2(obj, name, H5T_STD_I32LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return at.ok() && H5Awrite(at.id, H5T_NATIVE_INT32, v) >= 0;
}

// scalar fixed-length, null-padded ASCII string attribute of `len` bytes
// (h5py's dtype="S<len>")
bool attr_str(hid_t obj, const char* name, const char* s, size_t len) {
    std::vector<char> buf(len, '\0');
    std::memcpy(buf.data(), s, std::min(len, std::strlen(s)));
    H ty(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!ty.ok() || H5Tset_size(ty.id, len) < 0 || H5Tset_strpad(ty.id, H5T_STR_NULLPAD) < 0 ||
        H5Tset_cset(ty.id, H5T_CSET_ASCII) < 0)
        return false;
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    if (!sp.ok()) return false;
    H at(H5Acreate2(obj, name, ty.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return at.ok() && H5Awrite(at.id, ty.id, buf.data()) >= 0;
}

bool read_attr_ints(hid_t obj, const char* name, int32_t* v, int n) {
    if (H5Aexists(obj, name) <= 0) return false;
    H at(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    if (!at.ok()) return false;
    H sp(H5Aget_space(at.id), H5Sclose);
    if (!sp.ok() || H5Sget_simple_extent_npoints(sp.id) != n) return false;
    return H5Aread(at.id, H5T_NATIVE_INT32, v) >= 0;
}

bool valid_block(const osb_io_block* b) {
    if (!b || !b->name || b->ndim < 1 || b->ndim > 3 || b->halo < 0) return false;
    for (int d = 0; d < b->ndim; ++d)
        if (b->size[d] < 1) return false;
    return true;
}

}  // namespace

extern "C" {

const char* osb_io_last_error(void) { return g_io_error.c_str(); }

int osb_io_write(const char* path, const osb_io_block* b, int narrays, const char* const* names, const double* const* data,
                 int ndconst, const char* const* dnames, const double* dvals, int niconst, const char* const* inames,
                 const int32_t* ivals) {
    if (!path || !valid_block(b)) return fail(OSB_E_INVALID, "bad path or block description");
    if (narrays < 0 || (narrays > 0 && (!names || !data))) return fail(OSB_E_INVALID, "bad array list");
    if (ndconst < 0 || (ndconst > 0 && (!dnames || !dvals)) || niconst < 0 || (niconst > 0 && (!inames || !ivals)))
        return fail(OSB_E_INVALID, "bad constant list");
    std::lock_guard<std::mutex> lk(g_h5_mutex);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OSB_E_INVALID, std::string("cannot create ") + path);
    // ops_fetch_block_hdf5_file: the block group and its attributes
    H g(H5Gcreate2(f.id, b->name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(OSB_E_INVALID, std::string("cannot create group ") + b->name);
    const int32_t dims1 = b->ndim, index1 = b->index;
    if (!attr_ints(g.id, "dims", &dims1, 1) || !attr_str(g.id, "ops_type", "ops_block", 9) ||
        !attr_ints(g.id, "index", &index1, 1))
        return fail(OSB_E_INVALID, "cannot write block attributes");
    // ops_fetch_dat_hdf5_file: each array with its metadata
    const int nd = b->ndim;
    hsize_t shape[3];
    int32_t dm[3], dp[3], size[3], base[3];
    for (int d = 0; d < nd; ++d) {
        shape[nd - 1 - d] = (hsize_t)(b->size[d] + 2 * b->halo);   // slowest axis first
        dm[d] = -b->halo;
        dp[d] = b->halo;
        size[d] = b->size[d];
        base[d] = 0;
    }
    const int32_t one = 1;
    for (int a = 0; a < narrays; ++a) {
        if (!names[a] || !data[a]) return fail(OSB_E_INVALID, "null array name or data");
        H sp(H5Screate_simple(nd, shape, nullptr), H5Sclose);
        H ds(H5Dcreate2(g.id, names[a], H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return fail(OSB_E_INVALID, std::string("cannot create dataset ") + names[a]);
        if (H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data[a]) < 0)
            return fail(OSB_E_INVALID, std::string("cannot write dataset ") + names[a]);
        // set_hdf5_metadata (helperfunctions.py:99-116), same order
        if (!attr_ints(ds.id, "d_p", dp, nd) || !attr_ints(ds.id, "d_m", dm, nd) || !attr_ints(ds.id, "dim", &one, 1) ||
            !attr_str(ds.id, "ops_type", "ops_dat", 10) || !attr_ints(ds.id, "block_index", &index1, 1) ||
            !attr_ints(ds.id, "base", base, nd) || !attr_str(ds.id, "type", "double", 15) ||
            !attr_str(ds.id, "block", b->name, 25) || !attr_ints(ds.id, "size", size, nd))
            return fail(OSB_E_INVALID, std::string("cannot write attributes of ") + names[a]);
    }
    // write_constants (io_hdf5.py:165-179): one-element datasets at the root
    const hsize_t n1 = 1;
    for (int i = 0; i < ndconst + niconst; ++i) {
        const bool isint = i >= ndconst;
        const char* nm = isint ? inames[i - ndconst] : dnames[i];
        if (!nm) return fail(OSB_E_INVALID, "null constant name");
        H sp(H5Screate_simple(1, &n1, nullptr), H5Sclose);
        H ds(H5Dcreate2(f.id, nm, isint ? H5T_STD_I32LE : H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
             H5Dclose);
        if (!ds.ok()) return fail(OSB_E_INVALID, std::string("cannot create constant ") + nm);
        const herr_t e = isint ? H5Dwrite(ds.id, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, &ivals[i - ndconst])
                               : H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &dvals[i]);
        if (e < 0) return fail(OSB_E_INVALID, std::string("cannot write constant ") + nm);
    }
    if (H5Fflush(f.id, H5F_SCOPE_LOCAL) < 0) return fail(OSB_E_INVALID, "flush failed");
    return OSB_OK;
}

int osb_io_dat_info(const char* path, const char* block, const char* name, int32_t* ndim, int32_t size[3],
                    int32_t d_m[3], int32_t d_p[3]) {
    if (!path || !block || !name || !ndim || !size || !d_m || !d_p) return fail(OSB_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(g_h5_mutex);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OSB_E_INVALID, std::string("cannot open ") + path);
    const std::string full = std::string(block) + "/" + name;
    if (H5Lexists(f.id, block, H5P_DEFAULT) <= 0 || H5Lexists(f.id, full.c_str(), H5P_DEFAULT) <= 0)
        return fail(OSB_E_INVALID, "no dataset " + full + " in " + path);
    H ds(H5Dopen2(f.id, full.c_str(), H5P_DEFAULT), H5Dclose);
    H sp(H5Dget_space(ds.id), H5Sclose);
    const int rank = H5Sget_simple_extent_ndims(sp.id);
    if (rank < 1 || rank > 3) return fail(OSB_E_INVALID, "dataset rank must be 1..3");
    hsize_t shape[3] = {1, 1, 1};
    H5Sget_simple_extent_dims(sp.id, shape, nullptr);
    int32_t m[3] = {0, 0, 0}, p[3] = {0, 0, 0}, s[3] = {1, 1, 1};
    if (!read_attr_ints(ds.id, "d_m", m, rank) || !read_attr_ints(ds.id, "d_p", p, rank))
        return fail(OSB_E_INVALID, full + ": missing d_m/d_p attributes");
    for (int d = 0; d < rank; ++d) s[d] = (int32_t)shape[rank - 1 - d] + m[d] - p[d];
    int32_t sz[3] = {0, 0, 0};
    if (read_attr_ints(ds.id, "size", sz, rank)) {
        for (int d = 0; d < rank; ++d)
            if (sz[d] != s[d]) return fail(OSB_E_INVALID, full + ": size attribute disagrees with the dataset shape");
    }
    *ndim = rank;
    for (int d = 0; d < 3; ++d) {
        size[d] = d < rank ? s[d] : 1;
        d_m[d] = d < rank ? m[d] : 0;
        d_p[d] = d < rank ? p[d] : 0;
    }
    return OSB_OK;
}

int osb_io_read_dat(const char* path, const char* block, const char* name, double* out, int64_t count) {
    if (!path || !block || !name || !out) return fail(OSB_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(g_h5_mutex);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OSB_E_INVALID, std::string("cannot open ") + path);
    const std::string full = std::string(block) + "/" + name;
    if (H5Lexists(f.id, block, H5P_DEFAULT) <= 0 || H5Lexists(f.id, full.c_str(), H5P_DEFAULT) <= 0)
        return fail(OSB_E_INVALID, "no dataset " + full + " in " + path);
    H ds(H5Dopen2(f.id, full.c_str(), H5P_DEFAULT), H5Dclose);
    H sp(H5Dget_space(ds.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp.id) != count)
        return fail(OSB_E_INVALID, full + ": element count differs from the buffer");
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(OSB_E_INVALID, "cannot read " + full);
    return OSB_OK;
}

int osb_io_read_const(const char* path, const char* name, double* out) {
    if (!path || !name || !out) return fail(OSB_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(g_h5_mutex);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(OSB_E_INVALID, std::string("cannot open ") + path);
    if (H5Lexists(f.id, name, H5P_DEFAULT) <= 0) return fail(OSB_E_INVALID, std::string("no constant ") + name + " in " + path);
    H ds(H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose);
    H sp(H5Dget_space(ds.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(sp.id) < 1) return fail(OSB_E_INVALID, std::string("empty constant ") + name);
    std::vector<double> v((size_t)H5Sget_simple_extent_npoints(sp.id));
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
        return fail(OSB_E_INVALID, std::string("cannot read constant ") + name);
    *out = v[0];
    return OSB_OK;
}

}  // extern "C"
