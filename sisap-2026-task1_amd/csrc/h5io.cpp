// h5io.cpp -- HDF5 boundary of the kNN-graph builder (libknng_h5.so).
//
// The reference reads /train as NATIVE_FLOAT hyperslabs (src/main.cpp:147-177)
// and writes results/<task>/apex_simd.h5 with int32 `knns`, float32 `dists`
// and five root attributes (src/main.cpp:101-112, 459-480).  This file gives
// the same contract through the HDF5 C API, as plain C entry points that the
// CLI (knng_main) and the tests share.  Errors are reported as negative
// return codes with HDF5's own error stack silenced.

#include <hdf5.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "knng_h5.h"

namespace {

struct Quiet {   // silence HDF5's automatic error printing for one call
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct Ids {   // close-on-exit set of HDF5 handles
  hid_t file = -1, dset = -1, fspace = -1, mspace = -1, plist = -1;
  ~Ids() {
    if (plist >= 0) H5Pclose(plist);
    if (mspace >= 0) H5Sclose(mspace);
    if (fspace >= 0) H5Sclose(fspace);
    if (dset >= 0) H5Dclose(dset);
    if (file >= 0) H5Fclose(file);
  }
};

int read_block(const char* path, const char* name, int64_t row0, int64_t nrows, int64_t ncols,
               hid_t memtype, void* out) {
  Quiet q;
  Ids h;
  if ((h.file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) < 0) return KNNG_H5_EOPEN;
  if ((h.dset = H5Dopen2(h.file, name, H5P_DEFAULT)) < 0) return KNNG_H5_ENODSET;
  if ((h.fspace = H5Dget_space(h.dset)) < 0) return KNNG_H5_EIO;
  if (H5Sget_simple_extent_ndims(h.fspace) != 2) return KNNG_H5_ESHAPE;
  hsize_t dims[2];
  H5Sget_simple_extent_dims(h.fspace, dims, nullptr);
  if (row0 < 0 || nrows < 0 || hsize_t(row0 + nrows) > dims[0] || ncols < 0 || hsize_t(ncols) > dims[1])
    return KNNG_H5_ESHAPE;
  if (nrows == 0 || ncols == 0) return 0;
  hsize_t off[2] = {hsize_t(row0), 0}, cnt[2] = {hsize_t(nrows), hsize_t(ncols)};
  if (H5Sselect_hyperslab(h.fspace, H5S_SELECT_SET, off, nullptr, cnt, nullptr) < 0) return KNNG_H5_EIO;
  if ((h.mspace = H5Screate_simple(2, cnt, nullptr)) < 0) return KNNG_H5_EIO;
  if (H5Dread(h.dset, memtype, h.mspace, h.fspace, H5P_DEFAULT, out) < 0) return KNNG_H5_EIO;
  return 0;
}

int write_2d(hid_t file, const char* name, hid_t filetype, hid_t memtype, const void* data,
             int64_t rows, int64_t cols, int64_t chunk_rows) {
  Ids h;
  hsize_t dims[2] = {hsize_t(rows), hsize_t(cols)};
  if ((h.fspace = H5Screate_simple(2, dims, nullptr)) < 0) return KNNG_H5_EIO;
  h.plist = H5Pcreate(H5P_DATASET_CREATE);
  if (chunk_rows > 0 && rows > 0 && cols > 0) {
    hsize_t ch[2] = {hsize_t(chunk_rows < rows ? chunk_rows : rows), hsize_t(cols)};
    if (H5Pset_chunk(h.plist, 2, ch) < 0) return KNNG_H5_EIO;
  }
  if ((h.dset = H5Dcreate2(file, name, filetype, h.fspace, H5P_DEFAULT, h.plist, H5P_DEFAULT)) < 0)
    return KNNG_H5_EIO;
  if (rows > 0 && cols > 0 && H5Dwrite(h.dset, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return KNNG_H5_EIO;
  return 0;
}

int put_str_attr(hid_t file, const char* name, const char* value) {
  // variable-length string, scalar dataspace: src/main.cpp:101-106
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, H5T_VARIABLE);
  hid_t sp = H5Screate(H5S_SCALAR);
  int rc = 0;
  hid_t a = H5Acreate2(file, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0 || H5Awrite(a, t, &value) < 0) rc = KNNG_H5_EIO;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  H5Tclose(t);
  return rc;
}

int put_f32_attr(hid_t file, const char* name, float value) {
  // src/main.cpp:108-112
  hid_t sp = H5Screate(H5S_SCALAR);
  int rc = 0;
  hid_t a = H5Acreate2(file, name, H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0 || H5Awrite(a, H5T_NATIVE_FLOAT, &value) < 0) rc = KNNG_H5_EIO;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc;
}

}  // namespace

extern "C" {

struct knng_h5_reader {
  hid_t file = -1, dset = -1, fspace = -1;
  int64_t rows = 0, cols = 0;
};

int knng_h5_open(const char* path, const char* dset, knng_h5_reader** out) {
  if (!path || !dset || !out) return KNNG_H5_EINVAL;
  Quiet q;
  *out = nullptr;
  auto* r = new knng_h5_reader();
  if ((r->file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) < 0) {
    knng_h5_close(r);
    return KNNG_H5_EOPEN;
  }
  if ((r->dset = H5Dopen2(r->file, dset, H5P_DEFAULT)) < 0) {
    knng_h5_close(r);
    return KNNG_H5_ENODSET;
  }
  r->fspace = H5Dget_space(r->dset);
  if (r->fspace < 0 || H5Sget_simple_extent_ndims(r->fspace) != 2) {
    knng_h5_close(r);
    return KNNG_H5_ESHAPE;
  }
  hsize_t dims[2];
  H5Sget_simple_extent_dims(r->fspace, dims, nullptr);
  r->rows = int64_t(dims[0]);
  r->cols = int64_t(dims[1]);
  *out = r;
  return 0;
}

int knng_h5_shape_of(const knng_h5_reader* r, int64_t* rows, int64_t* cols) {
  if (!r) return KNNG_H5_EINVAL;
  if (rows) *rows = r->rows;
  if (cols) *cols = r->cols;
  return 0;
}

// Rows [row0, row0+nrows), columns [0, ncols) as NATIVE_FLOAT: the
// reference's hyperslab read, src/main.cpp:171-177.
int knng_h5_read_rows(knng_h5_reader* r, int64_t row0, int64_t nrows, int64_t ncols, float* out) {
  if (!r || !out) return KNNG_H5_EINVAL;
  if (row0 < 0 || nrows < 0 || row0 + nrows > r->rows || ncols < 0 || ncols > r->cols)
    return KNNG_H5_ESHAPE;
  if (nrows == 0 || ncols == 0) return 0;
  Quiet q;
  hsize_t off[2] = {hsize_t(row0), 0}, cnt[2] = {hsize_t(nrows), hsize_t(ncols)};
  if (H5Sselect_hyperslab(r->fspace, H5S_SELECT_SET, off, nullptr, cnt, nullptr) < 0) return KNNG_H5_EIO;
  hid_t ms = H5Screate_simple(2, cnt, nullptr);
  const herr_t e = H5Dread(r->dset, H5T_NATIVE_FLOAT, ms, r->fspace, H5P_DEFAULT, out);
  H5Sclose(ms);
  return e < 0 ? KNNG_H5_EIO : 0;
}

int knng_h5_close(knng_h5_reader* r) {
  if (!r) return 0;
  if (r->fspace >= 0) H5Sclose(r->fspace);
  if (r->dset >= 0) H5Dclose(r->dset);
  if (r->file >= 0) H5Fclose(r->file);
  delete r;
  return 0;
}

int knng_h5_shape(const char* path, const char* dset, int64_t* rows, int64_t* cols) {
  knng_h5_reader* r = nullptr;
  const int rc = knng_h5_open(path, dset, &r);
  if (rc) return rc;
  knng_h5_shape_of(r, rows, cols);
  knng_h5_close(r);
  return 0;
}

int knng_h5_read_f32(const char* path, const char* dset, int64_t row0, int64_t nrows, int64_t ncols,
                     float* out) {
  return read_block(path, dset, row0, nrows, ncols, H5T_NATIVE_FLOAT, out);
}

int knng_h5_read_i32(const char* path, const char* dset, int64_t row0, int64_t nrows, int64_t ncols,
                     int32_t* out) {
  return read_block(path, dset, row0, nrows, ncols, H5T_NATIVE_INT32, out);
}

int knng_h5_write_f32(const char* path, const char* dset, const float* data, int64_t rows,
                      int64_t cols, int64_t chunk_rows, int32_t append) {
  Quiet q;
  Ids h;
  h.file = append ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                  : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (h.file < 0) return KNNG_H5_EOPEN;
  return write_2d(h.file, dset, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, data, rows, cols, chunk_rows);
}

int knng_h5_write_i32(const char* path, const char* dset, const int32_t* data, int64_t rows,
                      int64_t cols, int64_t chunk_rows, int32_t append) {
  Quiet q;
  Ids h;
  h.file = append ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                  : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (h.file < 0) return KNNG_H5_EOPEN;
  return write_2d(h.file, dset, H5T_STD_I32LE, H5T_NATIVE_INT32, data, rows, cols, chunk_rows);
}

// Chunked writer for large /train files (rows appended in slices).
struct knng_h5_writer {
  hid_t file = -1, dset = -1;
  int64_t rows = 0, cols = 0;
};

int knng_h5_create_f32(const char* path, const char* dset, int64_t rows, int64_t cols,
                       int64_t chunk_rows, knng_h5_writer** out) {
  if (!path || !dset || !out || rows < 0 || cols <= 0) return KNNG_H5_EINVAL;
  Quiet q;
  auto* w = new knng_h5_writer();
  w->rows = rows;
  w->cols = cols;
  if ((w->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)) < 0) {
    delete w;
    return KNNG_H5_EOPEN;
  }
  hsize_t dims[2] = {hsize_t(rows), hsize_t(cols)};
  hid_t sp = H5Screate_simple(2, dims, nullptr);
  hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
  if (chunk_rows > 0 && rows > 0) {
    hsize_t ch[2] = {hsize_t(chunk_rows < rows ? chunk_rows : rows), hsize_t(cols)};
    H5Pset_chunk(pl, 2, ch);
  }
  w->dset = H5Dcreate2(w->file, dset, H5T_IEEE_F32LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
  H5Pclose(pl);
  H5Sclose(sp);
  if (w->dset < 0) {
    H5Fclose(w->file);
    delete w;
    return KNNG_H5_EIO;
  }
  *out = w;
  return 0;
}

int knng_h5_write_rows(knng_h5_writer* w, int64_t row0, int64_t nrows, const float* data) {
  if (!w || !data || row0 < 0 || nrows < 0 || row0 + nrows > w->rows) return KNNG_H5_EINVAL;
  if (nrows == 0) return 0;
  Quiet q;
  hid_t fs = H5Dget_space(w->dset);
  hsize_t off[2] = {hsize_t(row0), 0}, cnt[2] = {hsize_t(nrows), hsize_t(w->cols)};
  H5Sselect_hyperslab(fs, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
  hid_t ms = H5Screate_simple(2, cnt, nullptr);
  const herr_t e = H5Dwrite(w->dset, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, data);
  H5Sclose(ms);
  H5Sclose(fs);
  return e < 0 ? KNNG_H5_EIO : 0;
}

int knng_h5_finish(knng_h5_writer* w) {
  if (!w) return 0;
  if (w->dset >= 0) H5Dclose(w->dset);
  const herr_t e = w->file >= 0 ? H5Fclose(w->file) : 0;
  delete w;
  return e < 0 ? KNNG_H5_EIO : 0;
}

// The reference's output file: src/main.cpp:465-480.
int knng_h5_write_result(const char* path, const int32_t* knns, const float* dists, int64_t n,
                         int32_t k, const char* algo, const char* task, float buildtime,
                         float querytime, const char* params, const char* extra_name,
                         const char* extra_value) {
  if (!path || !knns || !dists || n < 0 || k <= 0) return KNNG_H5_EINVAL;
  Quiet q;
  Ids h;
  if ((h.file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)) < 0) return KNNG_H5_EOPEN;
  int rc;
  if ((rc = write_2d(h.file, "knns", H5T_NATIVE_INT, H5T_NATIVE_INT, knns, n, k, 0))) return rc;
  if ((rc = write_2d(h.file, "dists", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, dists, n, k, 0))) return rc;
  if ((rc = put_str_attr(h.file, "algo", algo ? algo : ""))) return rc;
  if ((rc = put_str_attr(h.file, "task", task ? task : ""))) return rc;
  if ((rc = put_f32_attr(h.file, "buildtime", buildtime))) return rc;
  if ((rc = put_f32_attr(h.file, "querytime", querytime))) return rc;
  if ((rc = put_str_attr(h.file, "params", params ? params : ""))) return rc;
  if (extra_name && extra_value && (rc = put_str_attr(h.file, extra_name, extra_value))) return rc;
  return 0;
}

int knng_h5_read_attr_str(const char* path, const char* name, char* buf, int64_t cap) {
  if (!path || !name || !buf || cap <= 0) return KNNG_H5_EINVAL;
  Quiet q;
  Ids h;
  if ((h.file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) < 0) return KNNG_H5_EOPEN;
  hid_t a = H5Aopen(h.file, name, H5P_DEFAULT);
  if (a < 0) return KNNG_H5_ENODSET;
  hid_t ft = H5Aget_type(a);
  int rc = 0;
  if (H5Tget_class(ft) != H5T_STRING) {
    rc = KNNG_H5_ESHAPE;
  } else if (H5Tis_variable_str(ft) > 0) {
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, H5T_VARIABLE);
    char* s = nullptr;
    if (H5Aread(a, mt, &s) < 0) {
      rc = KNNG_H5_EIO;
    } else {
      snprintf(buf, size_t(cap), "%s", s ? s : "");
      hid_t sp = H5Aget_space(a);
      H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, &s);
      H5Sclose(sp);
    }
    H5Tclose(mt);
  } else {
    const size_t len = H5Tget_size(ft);
    std::string tmp(len + 1, '\0');
    if (H5Aread(a, ft, &tmp[0]) < 0) rc = KNNG_H5_EIO;
    else snprintf(buf, size_t(cap), "%s", tmp.c_str());
  }
  H5Tclose(ft);
  H5Aclose(a);
  return rc;
}

int knng_h5_read_attr_f32(const char* path, const char* name, float* out) {
  if (!path || !name || !out) return KNNG_H5_EINVAL;
  Quiet q;
  Ids h;
  if ((h.file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) < 0) return KNNG_H5_EOPEN;
  hid_t a = H5Aopen(h.file, name, H5P_DEFAULT);
  if (a < 0) return KNNG_H5_ENODSET;
  const herr_t e = H5Aread(a, H5T_NATIVE_FLOAT, out);
  H5Aclose(a);
  return e < 0 ? KNNG_H5_EIO : 0;
}

}  // extern "C"
