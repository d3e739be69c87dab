/* HDF5 input/output of the reconstruction programs (include/trace_h5io.h).
 *
 * Plain C over the serial HDF5 1.10 C API, so that loading this library
 * pulls in libhdf5 and nothing else (no C++ runtime).  Follows the
 * reference's trace_io functions (src/common/trace_h5io.cc) for one process;
 * every HDF5 status is checked. */
#include "trace_h5io.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static _Thread_local char g_err[1024];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *trace_h5_last_error(void) { return g_err; }

/* HDF5 prints its error stack to stderr by default; the callers get a
 * status and a message instead. */
static void quiet(void) {
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  g_err[0] = '\0';
}

typedef struct {
  hid_t file, dset, space;
} Open;

static void close_open(Open *o) {
  if (o->space >= 0) H5Sclose(o->space);
  if (o->dset >= 0) H5Dclose(o->dset);
  if (o->file >= 0) H5Fclose(o->file);
  o->space = o->dset = o->file = -1;
}

/* Opens file/dataset read-only and checks that the dataset holds numbers
 * (ReadMetadata, trace_h5io.cc:34-45). */
static int open_dataset(const char *file, const char *dataset, Open *o) {
  o->file = o->dset = o->space = -1;
  if (!file || !dataset) return fail(TRACE_H5_E_INVALID, "null file or dataset path");
  quiet();
  o->file = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (o->file < 0) return fail(TRACE_H5_E_IO, "unable to open HDF5 file '%s'", file);
  o->dset = H5Dopen2(o->file, dataset, H5P_DEFAULT);
  if (o->dset < 0) {
    close_open(o);
    return fail(TRACE_H5_E_IO, "unable to open dataset '%s' in '%s'", dataset, file);
  }
  hid_t t = H5Dget_type(o->dset);
  H5T_class_t cls = t >= 0 ? H5Tget_class(t) : H5T_NO_CLASS;
  if (t >= 0) H5Tclose(t);
  if (cls != H5T_INTEGER && cls != H5T_FLOAT) {
    close_open(o);
    return fail(TRACE_H5_E_TYPE, "dataset '%s' in '%s' is not numeric", dataset, file);
  }
  o->space = H5Dget_space(o->dset);
  if (o->space < 0) {
    close_open(o);
    return fail(TRACE_H5_E_IO, "unable to get the dataspace of '%s'", dataset);
  }
  return TRACE_H5_OK;
}

int trace_h5_dims(const char *file, const char *dataset, int *ndims, int64_t *dims) {
  if (!ndims || !dims) return fail(TRACE_H5_E_INVALID, "null ndims or dims");
  Open o;
  int rc = open_dataset(file, dataset, &o);
  if (rc) return rc;
  const int nd = H5Sget_simple_extent_ndims(o.space);
  if (nd <= 0 || nd > TRACE_H5_MAX_DIMS) {
    close_open(&o);
    return fail(TRACE_H5_E_INVALID, "dataset '%s' has rank %d (1..%d supported)", dataset, nd,
                TRACE_H5_MAX_DIMS);
  }
  hsize_t d[TRACE_H5_MAX_DIMS];
  if (H5Sget_simple_extent_dims(o.space, d, NULL) != nd) {
    close_open(&o);
    return fail(TRACE_H5_E_IO, "unable to get the dimensions of '%s'", dataset);
  }
  *ndims = nd;
  for (int i = 0; i < nd; ++i) dims[i] = (int64_t)d[i];
  close_open(&o);
  return TRACE_H5_OK;
}

int trace_h5_read(const char *file, const char *dataset, float *out, int64_t count) {
  if (!out && count > 0) return fail(TRACE_H5_E_INVALID, "null output buffer");
  Open o;
  int rc = open_dataset(file, dataset, &o);
  if (rc) return rc;
  const hssize_t n = H5Sget_simple_extent_npoints(o.space);
  if (n != (hssize_t)count) {
    close_open(&o);
    return fail(TRACE_H5_E_INVALID, "dataset '%s' holds %lld values, the buffer %lld", dataset,
                (long long)n, (long long)count);
  }
  if (count > 0 && H5Dread(o.dset, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) {
    close_open(&o);
    return fail(TRACE_H5_E_IO, "unable to read dataset '%s' from '%s'", dataset, file);
  }
  close_open(&o);
  return TRACE_H5_OK;
}

int trace_h5_read_slices(const char *file, const char *dataset, int begin, int count,
                         float *out) {
  if (!out) return fail(TRACE_H5_E_INVALID, "null output buffer");
  Open o;
  int rc = open_dataset(file, dataset, &o);
  if (rc) return rc;
  hsize_t d[TRACE_H5_MAX_DIMS];
  const int nd = H5Sget_simple_extent_ndims(o.space);
  if (nd != 3) {  /* trace_h5io.cc:149-150 */
    close_open(&o);
    return fail(TRACE_H5_E_INVALID, "sinogram dataset '%s' must be 3-D [P][S][C], has rank %d",
                dataset, nd);
  }
  H5Sget_simple_extent_dims(o.space, d, NULL);
  if (begin < 0 || count < 1 || (hsize_t)begin + (hsize_t)count > d[1]) {
    close_open(&o);
    return fail(TRACE_H5_E_INVALID, "slices [%d, %d) outside the dataset's %llu slices", begin,
                begin + count, (unsigned long long)d[1]);
  }
  /* file: slices [begin, begin+count) of every projection (trace_h5io.cc:171-193);
   * memory: the whole [P][count][C] buffer */
  const hsize_t off[3] = {0, (hsize_t)begin, 0};
  const hsize_t cnt[3] = {d[0], (hsize_t)count, d[2]};
  hid_t mem = H5Screate_simple(3, cnt, NULL);
  herr_t st = mem >= 0 ? H5Sselect_hyperslab(o.space, H5S_SELECT_SET, off, NULL, cnt, NULL) : -1;
  if (st >= 0) st = H5Dread(o.dset, H5T_NATIVE_FLOAT, mem, o.space, H5P_DEFAULT, out);
  if (mem >= 0) H5Sclose(mem);
  close_open(&o);
  if (st < 0)
    return fail(TRACE_H5_E_IO, "unable to read slices [%d, %d) of '%s' from '%s'", begin,
                begin + count, dataset, file);
  return TRACE_H5_OK;
}

int trace_h5_write(const char *file, const char *dataset, int ndims, const int64_t *dims,
                   int type, const void *data, int truncate) {
  if (!file || !dataset || !dims) return fail(TRACE_H5_E_INVALID, "null argument");
  if (ndims < 1 || ndims > TRACE_H5_MAX_DIMS)
    return fail(TRACE_H5_E_INVALID, "rank %d outside 1..%d", ndims, TRACE_H5_MAX_DIMS);
  hsize_t d[TRACE_H5_MAX_DIMS];
  int64_t n = 1;
  for (int i = 0; i < ndims; ++i) {
    if (dims[i] < 0) return fail(TRACE_H5_E_INVALID, "negative extent in dimension %d", i);
    d[i] = (hsize_t)dims[i];
    n *= dims[i];
  }
  if (!data && n > 0) return fail(TRACE_H5_E_INVALID, "null data");
  hid_t mtype, ftype;
  switch (type) {
    case TRACE_H5_F32: mtype = H5T_NATIVE_FLOAT; ftype = H5T_IEEE_F32LE; break;
    case TRACE_H5_F64: mtype = H5T_NATIVE_DOUBLE; ftype = H5T_IEEE_F64LE; break;
    case TRACE_H5_U16: mtype = H5T_NATIVE_USHORT; ftype = H5T_STD_U16LE; break;
    case TRACE_H5_I32: mtype = H5T_NATIVE_INT; ftype = H5T_STD_I32LE; break;
    default: return fail(TRACE_H5_E_INVALID, "unknown element type %d", type);
  }
  quiet();
  hid_t f;
  if (truncate) {
    f = H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  } else {
    f = H5Fis_hdf5(file) > 0 ? H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT)
                            : H5Fcreate(file, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  }
  if (f < 0) return fail(TRACE_H5_E_IO, "unable to create or open '%s' for writing", file);
  hid_t space = H5Screate_simple(ndims, d, NULL);
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  if (lcpl >= 0) H5Pset_create_intermediate_group(lcpl, 1);
  hid_t ds = (space >= 0 && lcpl >= 0)
                 ? H5Dcreate2(f, dataset, ftype, space, lcpl, H5P_DEFAULT, H5P_DEFAULT)
                 : -1;
  int rc = TRACE_H5_OK;
  if (ds < 0) {
    rc = fail(TRACE_H5_E_IO, "unable to create dataset '%s' in '%s' (does it exist already?)",
              dataset, file);
  } else if (n > 0 && H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
    rc = fail(TRACE_H5_E_IO, "unable to write dataset '%s' to '%s'", dataset, file);
  }
  if (ds >= 0) H5Dclose(ds);
  if (lcpl >= 0) H5Pclose(lcpl);
  if (space >= 0) H5Sclose(space);
  if (H5Fclose(f) < 0 && rc == TRACE_H5_OK)
    rc = fail(TRACE_H5_E_IO, "unable to close '%s'", file);
  return rc;
}
