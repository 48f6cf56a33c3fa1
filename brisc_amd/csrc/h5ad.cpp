// .h5ad access for the PCA path (C ABI in include/bpca_h5ad.h): the X
// matrix, obs / var columns and obsm arrays as brisc's SingleCell reads and
// writes them (brisc/single_cell.py:83-1158 load, 3954 save; brisc/hdf5.pyx
// reads the same datasets), through the HDF5 C library.  Host code only; it
// is linked into its own library so that the PCA library does not depend on
// HDF5.  Reads go into caller buffers (pinned host memory when the caller
// streams them to the device), by slices of the first dimension, so a matrix
// larger than host memory can be streamed.
#include "bpca_h5ad.h"

#include <hdf5.h>

#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return -1;
}

// Closes an HDF5 identifier at scope exit.
struct Id {
    hid_t id;
    herr_t (*close)(hid_t);
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Id() {
        if (id >= 0) close(id);
    }
    operator hid_t() const { return id; }
    bool ok() const { return id >= 0; }
};

struct Quiet {  // no HDF5 error stack printing
    H5E_auto2_t func;
    void *data;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

// Element class codes of bpca_h5ad.h.
int class_code(hid_t type, int *size) {
    *size = (int)H5Tget_size(type);
    switch (H5Tget_class(type)) {
        case H5T_INTEGER: return H5Tget_sign(type) == H5T_SGN_NONE ? BH5_UINT : BH5_INT;
        case H5T_FLOAT: return BH5_FLOAT;
        case H5T_ENUM: return BH5_BOOL;   // numpy bool as written by h5py
        case H5T_STRING: return BH5_STRING;
        default: return -1;
    }
}

hid_t mem_type(int cls, int size) {
    if (cls == BH5_FLOAT) return size == 8 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
    if (cls == BH5_INT)
        return size == 8 ? H5T_NATIVE_INT64 : size == 4 ? H5T_NATIVE_INT32 : size == 2 ? H5T_NATIVE_INT16 : H5T_NATIVE_INT8;
    if (cls == BH5_UINT)
        return size == 8 ? H5T_NATIVE_UINT64 : size == 4 ? H5T_NATIVE_UINT32 : size == 2 ? H5T_NATIVE_UINT16 : H5T_NATIVE_UINT8;
    return -1;
}

// Creates the missing groups of `path`'s parent.
int make_parents(hid_t file, const std::string &path) {
    size_t pos = 0;
    while ((pos = path.find('/', pos + 1)) != std::string::npos) {
        const std::string g = path.substr(0, pos);
        if (g.empty()) continue;
        if (H5Lexists(file, g.c_str(), H5P_DEFAULT) > 0) continue;
        Id grp(H5Gcreate2(file, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!grp.ok()) return fail("cannot create group " + g);
    }
    return 0;
}

int set_string_attr(hid_t obj, const char *name, const char *value) {
    Quiet q;
    if (H5Aexists(obj, name) > 0) H5Adelete(obj, name);
    Id type(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(type, H5T_VARIABLE);
    H5Tset_cset(type, H5T_CSET_UTF8);
    Id space(H5Screate(H5S_SCALAR), H5Sclose);
    Id attr(H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("cannot write attribute ") + name);
    const char *v = value;
    return H5Awrite(attr, type, &v) < 0 ? fail(std::string("cannot write attribute ") + name) : 0;
}

}  // namespace

struct bh5_file {
    hid_t id = -1;
};

extern "C" {

const char *bh5_last_error(void) { return g_err.c_str(); }

int bh5_open(const char *path, int writable, bh5_file **out) {
    if (!path || !out) return fail("NULL argument");
    Quiet q;
    *out = nullptr;
    const hid_t id = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    if (id < 0) return fail(std::string("cannot open ") + path + " as an HDF5 file");
    auto *f = new bh5_file();
    f->id = id;
    *out = f;
    return 0;
}

int bh5_create(const char *path, bh5_file **out) {
    if (!path || !out) return fail("NULL argument");
    Quiet q;
    *out = nullptr;
    const hid_t id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (id < 0) return fail(std::string("cannot create ") + path);
    auto *f = new bh5_file();
    f->id = id;
    *out = f;
    return 0;
}

int bh5_close(bh5_file *f) {
    if (!f) return 0;
    const herr_t e = f->id >= 0 ? H5Fclose(f->id) : 0;
    delete f;
    return e < 0 ? fail("H5Fclose failed") : 0;
}

int bh5_kind(bh5_file *f, const char *path) {
    if (!f || !path) return fail("NULL argument");
    Quiet q;
    if (H5Lexists(f->id, path, H5P_DEFAULT) <= 0) return BH5_MISSING;
    H5O_info_t info;
    if (H5Oget_info_by_name(f->id, path, &info, H5P_DEFAULT) < 0) return fail(std::string("cannot inspect ") + path);
    return info.type == H5O_TYPE_GROUP ? BH5_GROUP : info.type == H5O_TYPE_DATASET ? BH5_DATASET : BH5_MISSING;
}

int bh5_dataset_info(bh5_file *f, const char *path, int *ndim, int64_t *dims,
                     int *cls, int *size) {
    if (!f || !path || !ndim || !dims || !cls || !size) return fail("NULL argument");
    Quiet q;
    Id ds(H5Dopen2(f->id, path, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("no dataset ") + path);
    Id type(H5Dget_type(ds), H5Tclose);
    Id space(H5Dget_space(ds), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(space);
    if (nd < 0 || nd > 2) return fail(std::string(path) + " has more than 2 dimensions");
    hsize_t d[2] = {1, 1};
    H5Sget_simple_extent_dims(space, d, nullptr);
    *ndim = nd;
    dims[0] = nd > 0 ? (int64_t)d[0] : 1;
    dims[1] = nd > 1 ? (int64_t)d[1] : 1;
    *cls = class_code(type, size);
    if (*cls < 0) return fail(std::string(path) + " has an unsupported element type");
    if (*cls == BH5_STRING && H5Tis_variable_str(type) > 0) *size = 0;
    return 0;
}

int bh5_read_rows(bh5_file *f, const char *path, int64_t first, int64_t count,
                  void *out) {
    if (!f || !path || (count > 0 && !out)) return fail("NULL argument");
    Quiet q;
    Id ds(H5Dopen2(f->id, path, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("no dataset ") + path);
    Id type(H5Dget_type(ds), H5Tclose);
    Id space(H5Dget_space(ds), H5Sclose);
    int size = 0;
    const int cls = class_code(type, &size);
    if (cls == BH5_STRING || cls < 0) return fail(std::string(path) + ": not a numeric dataset");
    const int nd = H5Sget_simple_extent_ndims(space);
    hsize_t d[2] = {1, 1};
    H5Sget_simple_extent_dims(space, d, nullptr);
    if (first < 0 || count < 0 || (nd > 0 && (hsize_t)(first + count) > d[0]))
        return fail(std::string(path) + ": rows out of range");
    if (count == 0) return 0;
    hsize_t start[2] = {(hsize_t)first, 0}, cnt[2] = {(hsize_t)count, d[1]};
    if (nd > 0 && H5Sselect_hyperslab(space, H5S_SELECT_SET, start, nullptr, cnt, nullptr) < 0)
        return fail("hyperslab selection failed");
    Id mspace(H5Screate_simple(nd > 0 ? nd : 1, cnt, nullptr), H5Sclose);
    // Booleans (enum) are read as their one-byte integer storage.
    hid_t mt = cls == BH5_BOOL ? H5T_NATIVE_UINT8 : mem_type(cls, size);
    Id mt_bool(cls == BH5_BOOL ? H5Tcopy(type) : -1, H5Tclose);
    if (cls == BH5_BOOL) mt = mt_bool;
    if (H5Dread(ds, mt, mspace, space, H5P_DEFAULT, out) < 0) return fail(std::string("cannot read ") + path);
    return 0;
}

int bh5_read_strings(bh5_file *f, const char *path, char *buffer,
                     int64_t buffer_size, int64_t *offsets, int64_t *total) {
    if (!f || !path || !total) return fail("NULL argument");
    Quiet q;
    Id ds(H5Dopen2(f->id, path, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("no dataset ") + path);
    Id type(H5Dget_type(ds), H5Tclose);
    if (H5Tget_class(type) != H5T_STRING) return fail(std::string(path) + ": not a string dataset");
    Id space(H5Dget_space(ds), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(space);
    std::vector<std::string> items((size_t)n);
    if (H5Tis_variable_str(type) > 0) {
        Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(type));
        std::vector<char *> ptrs((size_t)n, nullptr);
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptrs.data()) < 0)
            return fail(std::string("cannot read ") + path);
        for (hssize_t i = 0; i < n; i++) items[i] = ptrs[i] ? ptrs[i] : "";
        H5Dvlen_reclaim(mt, space, H5P_DEFAULT, ptrs.data());
    } else {
        const size_t w = H5Tget_size(type);
        std::vector<char> raw((size_t)n * w);
        if (H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw.data()) < 0)
            return fail(std::string("cannot read ") + path);
        for (hssize_t i = 0; i < n; i++) items[i] = std::string(raw.data() + i * w, strnlen(raw.data() + i * w, w));
    }
    int64_t tot = 0;
    for (auto &s : items) tot += (int64_t)s.size();
    *total = tot;
    if (!buffer || !offsets) return (int)n < 0 ? 0 : 0;  // size query
    if (buffer_size < tot) return fail("string buffer too small");
    int64_t at = 0;
    for (hssize_t i = 0; i < n; i++) {
        offsets[i] = at;
        std::memcpy(buffer + at, items[i].data(), items[i].size());
        at += (int64_t)items[i].size();
    }
    offsets[n] = at;
    return 0;
}

int bh5_attr_string(bh5_file *f, const char *path, const char *name, char *out,
                    int64_t out_size) {
    if (!f || !path || !name || !out || out_size < 1) return fail("NULL argument");
    Quiet q;
    out[0] = 0;
    if (H5Aexists_by_name(f->id, path, name, H5P_DEFAULT) <= 0) return 1;
    Id attr(H5Aopen_by_name(f->id, path, name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("cannot open attribute ") + name);
    Id type(H5Aget_type(attr), H5Tclose);
    if (H5Tget_class(type) != H5T_STRING) return fail(std::string("attribute ") + name + " is not a string");
    std::string v;
    if (H5Tis_variable_str(type) > 0) {
        Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(type));
        char *p = nullptr;
        if (H5Aread(attr, mt, &p) < 0) return fail("cannot read attribute");
        v = p ? p : "";
        Id sp(H5Screate(H5S_SCALAR), H5Sclose);
        H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, &p);
    } else {
        const size_t w = H5Tget_size(type);
        std::vector<char> raw(w + 1, 0);
        if (H5Aread(attr, type, raw.data()) < 0) return fail("cannot read attribute");
        v = std::string(raw.data(), strnlen(raw.data(), w));
    }
    if ((int64_t)v.size() + 1 > out_size) return fail("attribute buffer too small");
    std::memcpy(out, v.c_str(), v.size() + 1);
    return 0;
}

int bh5_attr_int64(bh5_file *f, const char *path, const char *name,
                   int64_t *out, int64_t max_count, int64_t *count) {
    if (!f || !path || !name || !out || !count) return fail("NULL argument");
    Quiet q;
    *count = 0;
    if (H5Aexists_by_name(f->id, path, name, H5P_DEFAULT) <= 0) return 1;
    Id attr(H5Aopen_by_name(f->id, path, name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    Id space(H5Aget_space(attr), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(space);
    if (n > max_count) return fail("attribute has too many values");
    if (H5Aread(attr, H5T_NATIVE_INT64, out) < 0) return fail(std::string("cannot read attribute ") + name);
    *count = n;
    return 0;
}

int bh5_list(bh5_file *f, const char *path, char *buffer, int64_t buffer_size,
             int64_t *needed) {
    if (!f || !path || !needed) return fail("NULL argument");
    Quiet q;
    Id grp(H5Gopen2(f->id, path, H5P_DEFAULT), H5Gclose);
    if (!grp.ok()) return fail(std::string("no group ") + path);
    H5G_info_t info;
    H5Gget_info(grp, &info);
    std::string names;
    for (hsize_t i = 0; i < info.nlinks; i++) {
        const ssize_t len = H5Lget_name_by_idx(grp, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
        if (len < 0) return fail("cannot list group");
        std::vector<char> nm((size_t)len + 1);
        H5Lget_name_by_idx(grp, ".", H5_INDEX_NAME, H5_ITER_INC, i, nm.data(), nm.size(), H5P_DEFAULT);
        names.append(nm.data());
        names.push_back('\n');
    }
    *needed = (int64_t)names.size() + 1;
    if (!buffer) return 0;
    if (buffer_size < *needed) return fail("name buffer too small");
    std::memcpy(buffer, names.c_str(), names.size() + 1);
    return 0;
}

int bh5_write_array(bh5_file *f, const char *path, const void *data, int ndim,
                    const int64_t *dims, int cls, int size,
                    const char *encoding_type) {
    if (!f || !path || !dims || ndim < 1 || ndim > 2) return fail("bad argument");
    Quiet q;
    if (make_parents(f->id, path) < 0) return -1;
    if (H5Lexists(f->id, path, H5P_DEFAULT) > 0 && H5Ldelete(f->id, path, H5P_DEFAULT) < 0)
        return fail(std::string("cannot replace ") + path);
    hsize_t d[2] = {(hsize_t)dims[0], ndim > 1 ? (hsize_t)dims[1] : 1};
    Id space(H5Screate_simple(ndim, d, nullptr), H5Sclose);
    hid_t mt = -1;
    Id bool_type(cls == BH5_BOOL ? H5Tenum_create(H5T_NATIVE_INT8) : -1, H5Tclose);
    if (cls == BH5_BOOL) {
        int8_t v0 = 0, v1 = 1;
        H5Tenum_insert(bool_type, "FALSE", &v0);
        H5Tenum_insert(bool_type, "TRUE", &v1);
        mt = bool_type;
    } else {
        mt = mem_type(cls, size);
    }
    if (mt < 0) return fail("unsupported element type");
    Id ds(H5Dcreate2(f->id, path, mt, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("cannot create ") + path);
    if (d[0] > 0 && H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(std::string("cannot write ") + path);
    if (encoding_type && encoding_type[0]) {
        if (set_string_attr(ds, "encoding-type", encoding_type) < 0) return -1;
        if (set_string_attr(ds, "encoding-version", "0.2.0") < 0) return -1;
    }
    return 0;
}

int bh5_write_strings(bh5_file *f, const char *path, const char *buffer,
                      const int64_t *offsets, int64_t count) {
    if (!f || !path || (count > 0 && (!buffer || !offsets))) return fail("bad argument");
    Quiet q;
    if (make_parents(f->id, path) < 0) return -1;
    if (H5Lexists(f->id, path, H5P_DEFAULT) > 0) H5Ldelete(f->id, path, H5P_DEFAULT);
    std::vector<std::string> items((size_t)count);
    std::vector<const char *> ptrs((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        items[i] = std::string(buffer + offsets[i], (size_t)(offsets[i + 1] - offsets[i]));
        ptrs[i] = items[i].c_str();
    }
    Id type(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(type, H5T_VARIABLE);
    H5Tset_cset(type, H5T_CSET_UTF8);
    hsize_t d[1] = {(hsize_t)count};
    Id space(H5Screate_simple(1, d, nullptr), H5Sclose);
    Id ds(H5Dcreate2(f->id, path, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("cannot create ") + path);
    if (count > 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptrs.data()) < 0)
        return fail(std::string("cannot write ") + path);
    if (set_string_attr(ds, "encoding-type", "string-array") < 0) return -1;
    return set_string_attr(ds, "encoding-version", "0.2.0");
}

int bh5_make_group(bh5_file *f, const char *path, const char *encoding_type) {
    if (!f || !path) return fail("NULL argument");
    Quiet q;
    std::string p(path);
    if (make_parents(f->id, p + "/x") < 0) return -1;
    if (encoding_type && encoding_type[0]) {
        Id grp(H5Gopen2(f->id, path, H5P_DEFAULT), H5Gclose);
        if (!grp.ok()) return fail(std::string("no group ") + path);
        if (set_string_attr(grp, "encoding-type", encoding_type) < 0) return -1;
        if (set_string_attr(grp, "encoding-version", std::string(encoding_type) == "dataframe" ? "0.2.0" : "0.1.0") < 0)
            return -1;
    }
    return 0;
}

int bh5_set_attr_string(bh5_file *f, const char *path, const char *name,
                        const char *value) {
    if (!f || !path || !name || !value) return fail("NULL argument");
    Quiet q;
    Id obj(H5Oopen(f->id, path, H5P_DEFAULT), H5Oclose);
    if (!obj.ok()) return fail(std::string("no object ") + path);
    return set_string_attr(obj, name, value);
}

int bh5_set_attr_strings(bh5_file *f, const char *path, const char *name,
                         const char *buffer, const int64_t *offsets, int64_t count) {
    if (!f || !path || !name) return fail("NULL argument");
    Quiet q;
    Id obj(H5Oopen(f->id, path, H5P_DEFAULT), H5Oclose);
    if (!obj.ok()) return fail(std::string("no object ") + path);
    if (H5Aexists(obj, name) > 0) H5Adelete(obj, name);
    std::vector<std::string> items((size_t)count);
    std::vector<const char *> ptrs((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        items[i] = std::string(buffer + offsets[i], (size_t)(offsets[i + 1] - offsets[i]));
        ptrs[i] = items[i].c_str();
    }
    Id type(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(type, H5T_VARIABLE);
    H5Tset_cset(type, H5T_CSET_UTF8);
    hsize_t d[1] = {(hsize_t)count};
    Id space(H5Screate_simple(1, d, nullptr), H5Sclose);
    Id attr(H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("cannot write attribute ") + name);
    return H5Awrite(attr, type, ptrs.data()) < 0 ? fail("cannot write attribute") : 0;
}

int bh5_set_attr_int64(bh5_file *f, const char *path, const char *name,
                       const int64_t *values, int64_t count) {
    if (!f || !path || !name || !values) return fail("NULL argument");
    Quiet q;
    Id obj(H5Oopen(f->id, path, H5P_DEFAULT), H5Oclose);
    if (!obj.ok()) return fail(std::string("no object ") + path);
    if (H5Aexists(obj, name) > 0) H5Adelete(obj, name);
    hsize_t d[1] = {(hsize_t)count};
    Id space(H5Screate_simple(1, d, nullptr), H5Sclose);
    Id attr(H5Acreate2(obj, name, H5T_STD_I64LE, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!attr.ok()) return fail(std::string("cannot write attribute ") + name);
    return H5Awrite(attr, H5T_NATIVE_INT64, values) < 0 ? fail("cannot write attribute") : 0;
}

}  // extern "C"
