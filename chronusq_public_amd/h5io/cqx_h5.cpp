// ChronusQ binary-file interop over the HDF5 C library (host code only).
//
// Mirrors the behaviour of ChronusQ's SafeFile (include/util/files.hpp):
// the same element types, the write-or-create rule of safeWriteData with
// nested group creation, and raw (column-major) matrix buffers.
#include "cqx_h5.h"

#include <hdf5.h>

#include <cstring>
#include <string>
#include <vector>

struct cqx_h5 {
  hid_t file = -1;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// HDF5 prints an error stack on every failed call unless told not to; the
// status codes carry the failure instead
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// complex<double> as the compound {double "r"; double "i"} (files.hpp:44-56)
hid_t complex_type() {
  hid_t t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE);
  return t;
}

// memory type of an element type; *owned says whether the caller closes it
hid_t mem_type(int dtype, bool *owned) {
  *owned = false;
  switch (dtype) {
    case CQX_H5_F64: return H5T_NATIVE_DOUBLE;
    case CQX_H5_U64: return H5T_NATIVE_UINT64;
    case CQX_H5_C128: *owned = true; return complex_type();
    default: return -1;
  }
}

// element type of a stored dataset type: F64 (float class, 8 bytes), U64
// (integer class, 8 bytes, unsigned), C128 (compound of two 8-byte floats
// named r and i); -1 otherwise
int classify(hid_t t) {
  const H5T_class_t c = H5Tget_class(t);
  const size_t sz = H5Tget_size(t);
  if (c == H5T_FLOAT && sz == 8) return CQX_H5_F64;
  if (c == H5T_INTEGER && sz == 8 && H5Tget_sign(t) == H5T_SGN_NONE) return CQX_H5_U64;
  if (c == H5T_COMPOUND && sz == 16 && H5Tget_nmembers(t) == 2) {
    bool ok = true;
    const char *want[2] = {"r", "i"};
    for (unsigned m = 0; m < 2; ++m) {
      char *nm = H5Tget_member_name(t, m);
      hid_t mt = H5Tget_member_type(t, m);
      ok = ok && nm && !std::strcmp(nm, want[m]) && H5Tget_class(mt) == H5T_FLOAT && H5Tget_size(mt) == 8 &&
           H5Tget_member_offset(t, m) == m * 8;
      if (nm) H5free_memory(nm);
      H5Tclose(mt);
    }
    if (ok) return CQX_H5_C128;
  }
  return -1;
}

// does every link of the path exist (checked one level at a time, as
// H5Lexists requires)
bool path_exists(hid_t file, const std::string &path) {
  size_t pos = 0;
  std::string cur;
  while (pos < path.size()) {
    size_t nx = path.find('/', pos);
    if (nx == std::string::npos) nx = path.size();
    const std::string tok = path.substr(pos, nx - pos);
    pos = nx + 1;
    if (tok.empty()) continue;
    cur += "/" + tok;
    if (H5Lexists(file, cur.c_str(), H5P_DEFAULT) <= 0) return false;
  }
  return !cur.empty();
}

// createGroup (files.hpp:79-101): every missing group along the path
bool make_groups(hid_t file, const std::string &group) {
  size_t pos = 0;
  std::string cur;
  while (pos < group.size()) {
    size_t nx = group.find('/', pos);
    if (nx == std::string::npos) nx = group.size();
    const std::string tok = group.substr(pos, nx - pos);
    pos = nx + 1;
    if (tok.empty()) continue;
    cur += "/" + tok;
    if (H5Lexists(file, cur.c_str(), H5P_DEFAULT) > 0) continue;
    hid_t g = H5Gcreate2(file, cur.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return false;
    H5Gclose(g);
  }
  return true;
}

std::string norm(const char *p) { return p[0] == '/' ? std::string(p) : "/" + std::string(p); }

}  // namespace

extern "C" {

const char *cqx_h5_last_error(void) { return g_err.c_str(); }

int cqx_h5_open(const char *path, int mode, cqx_h5 **out) {
  if (!path || !out || mode < CQX_H5_READ || mode > CQX_H5_CREATE) return fail(CQX_H5_ERR_ARG, "cqx_h5_open: bad arguments");
  quiet();
  hid_t f = mode == CQX_H5_CREATE ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                                  : H5Fopen(path, mode == CQX_H5_RDWR ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(CQX_H5_ERR_IO, std::string("cannot open ") + path);
  *out = new cqx_h5;
  (*out)->file = f;
  return 0;
}

int cqx_h5_close(cqx_h5 *f) {
  if (!f) return 0;
  const herr_t e = f->file >= 0 ? H5Fclose(f->file) : 0;
  delete f;
  return e < 0 ? fail(CQX_H5_ERR_IO, "cqx_h5_close: H5Fclose failed") : 0;
}

int cqx_h5_dims(cqx_h5 *f, const char *dataset, int *rank, unsigned long long *dims) {
  if (!f || !dataset || !rank || !dims) return fail(CQX_H5_ERR_ARG, "cqx_h5_dims: bad arguments");
  const std::string p = norm(dataset);
  if (!path_exists(f->file, p)) return fail(CQX_H5_ERR_MISSING, "no dataset " + p);
  hid_t d = H5Dopen2(f->file, p.c_str(), H5P_DEFAULT);
  if (d < 0) return fail(CQX_H5_ERR_MISSING, "no dataset " + p);
  hid_t s = H5Dget_space(d);
  const int r = H5Sget_simple_extent_ndims(s);
  int rc = 0;
  if (r < 0 || r > 8) {
    rc = fail(CQX_H5_ERR_IO, "unsupported dataspace of " + p);
  } else {
    hsize_t h[8];
    H5Sget_simple_extent_dims(s, h, nullptr);
    *rank = r;
    for (int i = 0; i < r; ++i) dims[i] = (unsigned long long)h[i];
  }
  H5Sclose(s);
  H5Dclose(d);
  return rc;
}

int cqx_h5_type(cqx_h5 *f, const char *dataset, int *dtype) {
  if (!f || !dataset || !dtype) return fail(CQX_H5_ERR_ARG, "cqx_h5_type: bad arguments");
  const std::string p = norm(dataset);
  if (!path_exists(f->file, p)) return fail(CQX_H5_ERR_MISSING, "no dataset " + p);
  hid_t d = H5Dopen2(f->file, p.c_str(), H5P_DEFAULT);
  if (d < 0) return fail(CQX_H5_ERR_MISSING, "no dataset " + p);
  hid_t t = H5Dget_type(d);
  const int c = classify(t);
  H5Tclose(t);
  H5Dclose(d);
  if (c < 0) return fail(CQX_H5_ERR_TYPE, "unsupported element type of " + p);
  *dtype = c;
  return 0;
}

int cqx_h5_read(cqx_h5 *f, const char *dataset, int dtype, void *buf, size_t count) {
  if (!f || !dataset || (!buf && count)) return fail(CQX_H5_ERR_ARG, "cqx_h5_read: bad arguments");
  int stored = -1;
  if (int rc = cqx_h5_type(f, dataset, &stored)) return rc;
  const std::string p = norm(dataset);
  if (stored != dtype) return fail(CQX_H5_ERR_TYPE, "element type of " + p + " differs from the requested one");
  hid_t d = H5Dopen2(f->file, p.c_str(), H5P_DEFAULT);
  hid_t s = H5Dget_space(d);
  const hssize_t n = H5Sget_simple_extent_npoints(s);
  H5Sclose(s);
  if (n < 0 || (size_t)n != count) {
    H5Dclose(d);
    return fail(CQX_H5_ERR_ARG, p + " holds " + std::to_string((long long)n) + " elements, not " +
                                    std::to_string(count));
  }
  bool owned = false;
  hid_t mt = mem_type(dtype, &owned);
  const herr_t e = count ? H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) : 0;
  if (owned) H5Tclose(mt);
  H5Dclose(d);
  return e < 0 ? fail(CQX_H5_ERR_IO, "H5Dread failed on " + p) : 0;
}

int cqx_h5_write(cqx_h5 *f, const char *dataset, int dtype, int rank, const unsigned long long *dims,
                 const void *buf) {
  if (!f || !dataset || rank < 1 || rank > 8 || !dims || !buf || dtype < CQX_H5_F64 || dtype > CQX_H5_U64)
    return fail(CQX_H5_ERR_ARG, "cqx_h5_write: bad arguments");
  const std::string p = norm(dataset);
  bool owned = false;
  hid_t mt = mem_type(dtype, &owned);
  hid_t d = -1;
  int rc = 0;
  if (path_exists(f->file, p)) {
    // writeData into the existing dataset: same extents, same type class
    d = H5Dopen2(f->file, p.c_str(), H5P_DEFAULT);
    int r = 0;
    unsigned long long have[8];
    int stored = -1;
    if (d < 0) {
      rc = fail(CQX_H5_ERR_IO, "cannot open " + p);
    } else if ((rc = cqx_h5_dims(f, dataset, &r, have)) == 0 && (rc = cqx_h5_type(f, dataset, &stored)) == 0) {
      bool same = r == rank && stored == dtype;
      for (int i = 0; same && i < rank; ++i) same = have[i] == dims[i];
      if (!same) rc = fail(CQX_H5_ERR_ARG, "existing dataset " + p + " has other extents or type");
    }
  } else {
    const size_t cut = p.rfind('/');
    if (cut > 0 && !make_groups(f->file, p.substr(0, cut))) {
      rc = fail(CQX_H5_ERR_IO, "cannot create the groups of " + p);
    } else {
      // file type: little-endian IEEE doubles / U64, or the {r, i} compound
      hid_t ft = dtype == CQX_H5_F64 ? H5Tcopy(H5T_IEEE_F64LE) : dtype == CQX_H5_U64 ? H5Tcopy(H5T_STD_U64LE)
                                                                                   : complex_type();
      std::vector<hsize_t> h(dims, dims + rank);
      hid_t s = H5Screate_simple(rank, h.data(), nullptr);
      d = H5Dcreate2(f->file, p.c_str(), ft, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      H5Sclose(s);
      H5Tclose(ft);
      if (d < 0) rc = fail(CQX_H5_ERR_IO, "cannot create " + p);
    }
  }
  if (rc == 0 && H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) rc = fail(CQX_H5_ERR_IO, "H5Dwrite failed on " + p);
  if (d >= 0) H5Dclose(d);
  if (owned) H5Tclose(mt);
  return rc;
}

}  // extern "C"
