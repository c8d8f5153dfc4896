// hdf5_io.h -- HMatrix / HVector in HDF5 files, the reference's layout
// (libs/libhmat/HDF5IO.cc); HDF5 itself is loaded at run time.
#pragma once

#include <string>

namespace scuffx {

// every function returns an empty string on success, else the error
std::string hdf5_open(const char *path, void **ctx);   // create / truncate
std::string hdf5_close(void *ctx);
// column-major host data; complex: interleaved re/im
std::string hdf5_write_matrix(void *ctx, const char *name, const double *data, int nr, int nc, int cplx);
std::string hdf5_write_vector(void *ctx, const char *name, const double *data, int n, int cplx);
// data may be null: sizes and type only
std::string hdf5_read_matrix(const char *path, const char *name, double *data, int *nr, int *nc, int *cplx);
std::string hdf5_read_vector(const char *path, const char *name, double *data, int *n, int *cplx);

}  // namespace scuffx
