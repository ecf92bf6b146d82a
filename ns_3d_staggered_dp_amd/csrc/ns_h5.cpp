// The reference's HDF5 file layout from C (include/ns_h5.h).  Host code
// only, linked against the HDF5 C and high-level libraries.
#include "ns_h5.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstring>
#include <string>
#include <vector>

namespace {

struct Quiet {
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

void quiet() { static Quiet q; }

// split "a/b/c" into parent "a/b" and leaf "c"; parent "" means the root
void split(const std::string &p, std::string &parent, std::string &leaf) {
    const size_t k = p.find_last_of('/');
    if (k == std::string::npos) {
        parent.clear();
        leaf = p;
    } else {
        parent = p.substr(0, k);
        leaf = p.substr(k + 1);
    }
}

// open (creating when asked) every group along path; returns the last one
hid_t open_groups(hid_t f, const std::string &path, bool create) {
    hid_t cur = H5Gopen2(f, "/", H5P_DEFAULT);
    if (cur < 0) return -1;
    size_t s = 0;
    while (s < path.size()) {
        size_t e = path.find('/', s);
        if (e == std::string::npos) e = path.size();
        const std::string name = path.substr(s, e - s);
        s = e + 1;
        if (name.empty()) continue;
        hid_t nxt;
        if (H5Lexists(cur, name.c_str(), H5P_DEFAULT) > 0) nxt = H5Gopen2(cur, name.c_str(), H5P_DEFAULT);
        else if (create) nxt = H5Gcreate2(cur, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        else nxt = -1;
        H5Gclose(cur);
        if (nxt < 0) return -1;
        cur = nxt;
    }
    return cur;
}

// C-order (HDF5) dims of a Fortran-order shape
void c_dims(int rank, const long *fd, hsize_t *cd) {
    for (int i = 0; i < rank; i++) cd[i] = (hsize_t)fd[rank - 1 - i];
}

}  // namespace

extern "C" {

int nsh5_create(const char *path, nsh5_file *f) {
    quiet();
    if (!path || !f) return NSH5_ERR_INVALID;
    hid_t id = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    if (id < 0) return NSH5_ERR_IO;
    *f = (nsh5_file)id;
    return NSH5_OK;
}

int nsh5_open(const char *path, int readwrite, nsh5_file *f) {
    quiet();
    if (!path || !f) return NSH5_ERR_INVALID;
    hid_t id = H5Fopen(path, readwrite ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    if (id < 0) return NSH5_ERR_IO;
    *f = (nsh5_file)id;
    return NSH5_OK;
}

int nsh5_close(nsh5_file f) { return H5Fclose((hid_t)f) < 0 ? NSH5_ERR_IO : NSH5_OK; }

int nsh5_exists(nsh5_file f, const char *path) {
    quiet();
    if (!path) return NSH5_ERR_INVALID;
    // every component must exist
    std::string p(path), acc;
    size_t s = 0;
    while (s < p.size()) {
        size_t e = p.find('/', s);
        if (e == std::string::npos) e = p.size();
        const std::string name = p.substr(s, e - s);
        s = e + 1;
        if (name.empty()) continue;
        acc += (acc.empty() ? "" : "/") + name;
        if (H5Lexists((hid_t)f, acc.c_str(), H5P_DEFAULT) <= 0) return 0;
    }
    return 1;
}

int nsh5_write_double(nsh5_file f, const char *path, int rank, const long *dims, const double *data) {
    quiet();
    if (!path || !dims || !data || rank < 1 || rank > 8) return NSH5_ERR_INVALID;
    std::string parent, leaf;
    split(path, parent, leaf);
    hid_t g = open_groups((hid_t)f, parent, true);
    if (g < 0) return NSH5_ERR_IO;
    hsize_t cd[8];
    c_dims(rank, dims, cd);
    herr_t e = H5LTmake_dataset_double(g, leaf.c_str(), rank, cd, data);
    H5Gclose(g);
    return e < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_get_dims(nsh5_file f, const char *path, int *rank, long *dims) {
    quiet();
    if (!path || !rank || !dims) return NSH5_ERR_INVALID;
    int r = 0;
    if (H5LTget_dataset_ndims((hid_t)f, path, &r) < 0 || r > 8) return NSH5_ERR_IO;
    hsize_t cd[8];
    H5T_class_t cls;
    size_t sz;
    if (H5LTget_dataset_info((hid_t)f, path, cd, &cls, &sz) < 0) return NSH5_ERR_IO;
    *rank = r;
    for (int i = 0; i < r; i++) dims[i] = (long)cd[r - 1 - i];
    return NSH5_OK;
}

int nsh5_read_double(nsh5_file f, const char *path, int rank, const long *dims, double *data) {
    if (!dims || !data) return NSH5_ERR_INVALID;
    int r = 0;
    long fd[8];
    int rc = nsh5_get_dims(f, path, &r, fd);
    if (rc) return rc;
    if (r != rank) return NSH5_ERR_SHAPE;
    for (int i = 0; i < r; i++)
        if (fd[i] != dims[i]) return NSH5_ERR_SHAPE;
    return H5LTread_dataset_double((hid_t)f, path, data) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_set_attr_int(nsh5_file f, const char *obj, const char *name, int v) {
    quiet();
    if (!obj || !name) return NSH5_ERR_INVALID;
    return H5LTset_attribute_int((hid_t)f, obj, name, &v, 1) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_set_attr_double(nsh5_file f, const char *obj, const char *name, double v) {
    quiet();
    if (!obj || !name) return NSH5_ERR_INVALID;
    return H5LTset_attribute_double((hid_t)f, obj, name, &v, 1) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_set_attr_string(nsh5_file f, const char *obj, const char *name, const char *v) {
    quiet();
    if (!obj || !name || !v) return NSH5_ERR_INVALID;
    return H5LTset_attribute_string((hid_t)f, obj, name, v) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_get_attr_int(nsh5_file f, const char *obj, const char *name, int *v) {
    quiet();
    if (!obj || !name || !v) return NSH5_ERR_INVALID;
    return H5LTget_attribute_int((hid_t)f, obj, name, v) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_get_attr_double(nsh5_file f, const char *obj, const char *name, double *v) {
    quiet();
    if (!obj || !name || !v) return NSH5_ERR_INVALID;
    return H5LTget_attribute_double((hid_t)f, obj, name, v) < 0 ? NSH5_ERR_IO : NSH5_OK;
}

int nsh5_get_attr_string(nsh5_file f, const char *obj, const char *name, char *buf, long cap) {
    quiet();
    if (!obj || !name || !buf || cap < 1) return NSH5_ERR_INVALID;
    // fixed-length strings of any padding (the Fortran API writes the trimmed length)
    hid_t a = H5Aopen_by_name((hid_t)f, obj, name, H5P_DEFAULT, H5P_DEFAULT);
    if (a < 0) return NSH5_ERR_IO;
    hid_t t = H5Aget_type(a);
    int rc = NSH5_ERR_IO;
    if (t >= 0 && H5Tget_class(t) == H5T_STRING && !H5Tis_variable_str(t)) {
        const size_t sz = H5Tget_size(t);
        std::vector<char> tmp(sz + 1, 0);
        if (H5Aread(a, t, tmp.data()) >= 0) {
            size_t n = strnlen(tmp.data(), sz);
            while (n > 0 && tmp[n - 1] == ' ') n--;
            if ((long)n < cap) {
                memcpy(buf, tmp.data(), n);
                buf[n] = 0;
                rc = NSH5_OK;
            } else {
                rc = NSH5_ERR_INVALID;
            }
        }
    }
    if (t >= 0) H5Tclose(t);
    H5Aclose(a);
    return rc;
}

int nsh5_list(nsh5_file f, const char *group, char *buf, long cap) {
    quiet();
    if (!group || !buf || cap < 1) return NSH5_ERR_INVALID;
    hid_t g = open_groups((hid_t)f, group, false);
    if (g < 0) return NSH5_ERR_IO;
    std::string out;
    hsize_t n = 0;
    H5G_info_t info;
    if (H5Gget_info(g, &info) >= 0) n = info.nlinks;
    for (hsize_t i = 0; i < n; i++) {
        ssize_t len = H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
        if (len < 0) continue;
        std::vector<char> nm(len + 1, 0);
        H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nm.data(), len + 1, H5P_DEFAULT);
        if (!out.empty()) out += '\n';
        out += nm.data();
    }
    H5Gclose(g);
    if ((long)out.size() >= cap) return NSH5_ERR_INVALID;
    memcpy(buf, out.c_str(), out.size() + 1);
    return NSH5_OK;
}

}  // extern "C"
