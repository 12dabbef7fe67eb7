/*
 * xpcs_h5.h -- HDF5 access of the drop-in host layer (libxpcs_host.so).
 *
 * The reference writes its results with H5Result (src/xpcs/h5_result.cpp:56-347)
 * and reads its configuration with Configuration (src/xpcs/configuration.cpp:386-552).
 * These entry points give the same datasets to C / Python callers without an
 * HDF5 binding of their own: open or create a file, write or replace a dataset
 * (parent groups are created), read a dataset converted to a native type.
 * libhdf5 is loaded at run time (XPCS_HDF5_LIB overrides the search).
 * Every function returns 0 on success; xpcs_h5_last_error() explains failures.
 */
#ifndef XPCS_H5_H
#define XPCS_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xpcs_h5_file xpcs_h5_file;

const char* xpcs_h5_last_error(void);
int32_t xpcs_h5_open(const char* path, int32_t create, xpcs_h5_file** f);
int32_t xpcs_h5_close(xpcs_h5_file* f);
/* dims[0..*ndims) of a dataset (ndims 0 for a scalar); error if missing. */
int32_t xpcs_h5_dims(xpcs_h5_file* f, const char* path, uint64_t* dims, int32_t cap, int32_t* ndims);
/* Stored element type of a dataset, as H5Result leaves it for a reader
 * (src/xpcs/h5_result.cpp:56-347 writes float32, or float64 for its double
 * overload): klass 0 integer, 1 float, 3 string (HDF5's H5T_class_t); size in
 * bytes; is_signed 1/0 for integers (-1 otherwise). */
int32_t xpcs_h5_dtype(xpcs_h5_file* f, const char* path, int32_t* klass, int32_t* size, int32_t* is_signed);
int32_t xpcs_h5_read_f64(xpcs_h5_file* f, const char* path, double* out, int64_t cap);
int32_t xpcs_h5_read_f32(xpcs_h5_file* f, const char* path, float* out, int64_t cap);
int32_t xpcs_h5_read_i32(xpcs_h5_file* f, const char* path, int32_t* out, int64_t cap);
/* string dataset (fixed or variable length) into buf (cap bytes incl. NUL) */
int32_t xpcs_h5_read_string(xpcs_h5_file* f, const char* path, char* buf, int64_t cap);
/* writers: dtype 0 = int32, 1 = int64, 2 = float32, 3 = float64, 4 = uint16, 5 = uint32 */
int32_t xpcs_h5_write(xpcs_h5_file* f, const char* path, int32_t dtype, const void* data,
                      const uint64_t* dims, int32_t ndims);
int32_t xpcs_h5_write_string(xpcs_h5_file* f, const char* path, const char* s);

#ifdef __cplusplus
}
#endif
#endif
