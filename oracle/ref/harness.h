// harness.h -- drives the reference's own hot-path classes (compiled from the
// reference sources by oracle/Makefile into oracle/_ref/) so that tests can
// compare the oracle restatement with the real reference.  Test
// infrastructure only.  Each ref_* entry mirrors the body of one reference
// main between input reading and output writing (e.g. sirt_main.cc:102-202),
// with the HDF5 reads replaced by caller buffers and MPI absent (one rank).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "trace_h5io.h"

namespace ref_harness {

// theta as trace_io::ReadTheta returns it: malloc'ed data (DegreeToRadian
// free()s it, trace_utils.cc:22) and a metadata record with dims[0] = P.
struct ThetaBox {
  trace_io::H5Metadata md;
  hsize_t dims[1];
  trace_io::H5Data data;
  ThetaBox(const float *theta, int P) {
    dims[0] = (hsize_t)P;
    md.ndims = 1;
    md.dims = dims;
    md.data_size = (hsize_t)P;
    data.data = std::malloc(sizeof(float) * (size_t)P);
    std::memcpy(data.data, theta, sizeof(float) * (size_t)P);
    data.offset = 0;
    data.count = (hsize_t)P;
    data.in_memory_type_size = sizeof(float);
    data.metadata = &md;
  }
  // ownership of data.data passes to TraceMetadata (delete[]), as in the mains
  float *release_new_array() {
    // TraceMetadata deletes theta with delete[] (trace_data.h:129); after
    // DegreeToRadian the buffer was allocated with new[] (trace_utils.cc:13).
    // Without DegreeToRadian (PML) copy into a new[] buffer.
    return static_cast<float *>(data.data);
  }
};

inline float *new_copy(const float *src, size_t n) {
  float *dst = new float[n];
  std::memcpy(dst, src, n * sizeof(float));
  return dst;
}

}  // namespace ref_harness
