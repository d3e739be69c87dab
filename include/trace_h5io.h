/*
 * trace_h5io.h - HDF5 input/output of the reconstruction programs.
 *
 * A C-ABI over the serial HDF5 C library (1.10).  It replaces the
 * reference's trace_io namespace (src/common/trace_h5io.h:8-64,
 * trace_h5io.cc) for one process:
 *
 *   reference                                      here
 *   trace_io::ReadMetadata   (trace_h5io.cc:24-83)   trace_h5_dims
 *   trace_io::ReadTheta      (trace_h5io.cc:85-133)  trace_h5_read
 *   trace_io::ReadSlices     (trace_h5io.cc:135-266) trace_h5_read_slices
 *   trace_io::WriteData / WriteRecon
 *                            (trace_h5io.cc:366-503) trace_h5_write
 *
 * Differences from the reference, all in the direction of refusing input
 * the reference would misread:
 *   - values are converted to float by HDF5 whatever the stored type is
 *     (the reference reads the file's native type into a buffer it then
 *     treats as float, so a float64 theta or an integer sinogram is
 *     reinterpreted bytes there);
 *   - every HDF5 status is checked and reported as a return code with the
 *     text in trace_h5_last_error() (the reference ignores them);
 *   - trace_h5_write creates missing intermediate groups ("/exchange" for
 *     "/exchange/data").
 *
 * Thread safety: the HDF5 library is built without thread safety; call
 * these functions from one thread at a time.
 */
#ifndef TRACE_H5IO_H_
#define TRACE_H5IO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRACE_H5_OK 0
#define TRACE_H5_E_INVALID (-1)  /* bad argument (null, rank, range) */
#define TRACE_H5_E_IO (-2)       /* HDF5 could not open, read or write */
#define TRACE_H5_E_TYPE (-3)     /* dataset is not of a numeric class */

#define TRACE_H5_MAX_DIMS 8

/* Element types for trace_h5_write. */
#define TRACE_H5_F32 0
#define TRACE_H5_F64 1
#define TRACE_H5_U16 2
#define TRACE_H5_I32 3

/* Rank and extent of a dataset (ReadMetadata, trace_h5io.cc:24-83). */
int trace_h5_dims(const char *file, const char *dataset, int *ndims, int64_t *dims);

/* The whole dataset as float, `count` elements (must equal the dataset's
 * element count).  Used for theta (ReadTheta, trace_h5io.cc:85-133). */
int trace_h5_read(const char *file, const char *dataset, float *out, int64_t count);

/* Slices [begin, begin+count) of a 3-D sinogram [P][S][C] as float into
 * out[P][count][C] (ReadSlices, trace_h5io.cc:135-266: hyperslab offset
 * {0, begin, 0}, block {P, 1, C}). */
int trace_h5_read_slices(const char *file, const char *dataset, int begin, int count,
                         float *out);

/* Writes an ndims-dimensional dataset of element type `type` (TRACE_H5_*).
 * truncate = 1 creates or truncates the file (H5F_ACC_TRUNC, as WriteData,
 * trace_h5io.cc:382); 0 opens an existing file read-write, or creates it,
 * and adds the dataset (which must not exist yet). */
int trace_h5_write(const char *file, const char *dataset, int ndims, const int64_t *dims,
                   int type, const void *data, int truncate);

/* Text of the last failure of this library (empty when none). */
const char *trace_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TRACE_H5IO_H_ */
