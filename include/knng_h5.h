/* knng_h5.h -- HDF5 boundary of the kNN-graph builder (libknng_h5.so).
 *
 * Replaces the reference's HDF5 C++ calls: the /train hyperslab ingest of
 * src/main.cpp:147-177 and the result file of src/main.cpp:101-112, 459-480
 * (int32 `knns`, float32 `dists`, root attributes algo/task/buildtime/
 * querytime/params).  All functions return 0 or a negative KNNG_H5_* code.
 */
#ifndef KNNG_H5_H
#define KNNG_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KNNG_H5_EINVAL -1
#define KNNG_H5_EOPEN -2    /* file missing or not HDF5 */
#define KNNG_H5_ENODSET -3  /* dataset or attribute missing */
#define KNNG_H5_ESHAPE -4   /* rank/extent/selection mismatch */
#define KNNG_H5_EIO -5      /* HDF5 read/write failure */

typedef struct knng_h5_reader knng_h5_reader;
typedef struct knng_h5_writer knng_h5_writer;

/* streaming reader of a rank-2 dataset (src/main.cpp:149-154) */
int knng_h5_open(const char* path, const char* dset, knng_h5_reader** out);
int knng_h5_shape_of(const knng_h5_reader* r, int64_t* rows, int64_t* cols);
/* rows [row0,row0+nrows) x columns [0,ncols) as NATIVE_FLOAT (src/main.cpp:171-177) */
int knng_h5_read_rows(knng_h5_reader* r, int64_t row0, int64_t nrows, int64_t ncols, float* out);
int knng_h5_close(knng_h5_reader* r);

/* one-shot helpers */
int knng_h5_shape(const char* path, const char* dset, int64_t* rows, int64_t* cols);
int knng_h5_read_f32(const char* path, const char* dset, int64_t row0, int64_t nrows, int64_t ncols,
                     float* out);
int knng_h5_read_i32(const char* path, const char* dset, int64_t row0, int64_t nrows, int64_t ncols,
                     int32_t* out);
int knng_h5_write_f32(const char* path, const char* dset, const float* data, int64_t rows,
                      int64_t cols, int64_t chunk_rows, int32_t append);
int knng_h5_write_i32(const char* path, const char* dset, const int32_t* data, int64_t rows,
                      int64_t cols, int64_t chunk_rows, int32_t append);

/* chunked float32 writer for large synthetic /train files */
int knng_h5_create_f32(const char* path, const char* dset, int64_t rows, int64_t cols,
                       int64_t chunk_rows, knng_h5_writer** out);
int knng_h5_write_rows(knng_h5_writer* w, int64_t row0, int64_t nrows, const float* data);
int knng_h5_finish(knng_h5_writer* w);

/* the reference's result file (src/main.cpp:465-480); extra_name/extra_value
 * add one more string attribute when both are non-NULL */
int knng_h5_write_result(const char* path, const int32_t* knns, const float* dists, int64_t n,
                         int32_t k, const char* algo, const char* task, float buildtime,
                         float querytime, const char* params, const char* extra_name,
                         const char* extra_value);
int knng_h5_read_attr_str(const char* path, const char* name, char* buf, int64_t cap);
int knng_h5_read_attr_f32(const char* path, const char* name, float* out);

#ifdef __cplusplus
}
#endif
#endif /* KNNG_H5_H */
