// hdf5_io.cpp -- HMatrix / HVector export to and import from HDF5 files in
// the reference's layout (libs/libhmat/HDF5IO.cc:120-540).
//
// The HDF5 library is loaded at run time (libhdf5 + its high-level lite
// API libhdf5_hl: from the loader path, $SCUFFX_HDF5_DIR, or /opt/conda/lib),
// so libscuffx has no link-time dependency on it.  Layout:
//   matrix (normal storage), column-major data: dataset dims {NC, NR} (real)
//   or {NC, 2 NR} (complex, interleaved re/im), int attributes
//   "Storage_Type" = 0 and "RealComplex" = 0 / 1;
//   vector: dims {N} or {2 N}, int attribute "RealComplex".
// HDF5 is row-major, so a reader sees the transpose's shape, as with the
// reference's files.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <string>

#include "hdf5_io.h"

namespace scuffx {

namespace {

typedef int64_t hid_t;   // HDF5 >= 1.10
typedef int herr_t;
typedef unsigned long long hsize_t;
const hid_t kDefault = 0;          // H5P_DEFAULT
const unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;

struct H5 {
  bool ok = false;
  std::string err;
  herr_t (*open)();
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  herr_t (*eset_auto2)(hid_t, void *, void *);
  herr_t (*make_double)(hid_t, const char *, int, const hsize_t *, const double *);
  herr_t (*read_double)(hid_t, const char *, double *);
  herr_t (*set_attr_int)(hid_t, const char *, const char *, const int *, size_t);
  herr_t (*get_attr_int)(hid_t, const char *, const char *, int *);
  herr_t (*find_dataset)(hid_t, const char *);
  herr_t (*get_ndims)(hid_t, const char *, int *);
  herr_t (*get_info)(hid_t, const char *, hsize_t *, int *, size_t *);
};

void *open_lib(const char *name)
{
  if (void *h = dlopen(name, RTLD_NOW | RTLD_GLOBAL)) return h;
  std::string dirs[2] = {getenv("SCUFFX_HDF5_DIR") ? getenv("SCUFFX_HDF5_DIR") : "", "/opt/conda/lib"};
  for (const std::string &d : dirs)
    if (!d.empty())
      if (void *h = dlopen((d + "/" + name).c_str(), RTLD_NOW | RTLD_GLOBAL)) return h;
  return nullptr;
}

template <typename F>
bool bind(void *lib, const char *sym, F *f)
{
  *f = (F)dlsym(lib, sym);
  return *f != nullptr;
}

H5 &h5()
{
  static H5 L;
  static bool tried = false;
  if (tried) return L;
  tried = true;
  void *a = open_lib("libhdf5.so"), *b = a ? open_lib("libhdf5_hl.so") : nullptr;
  if (!a || !b) { L.err = "HDF5 libraries (libhdf5.so, libhdf5_hl.so) not found"; return L; }
  bool ok = bind(a, "H5open", &L.open) && bind(a, "H5Fcreate", &L.fcreate) && bind(a, "H5Fopen", &L.fopen) &&
            bind(a, "H5Fclose", &L.fclose) && bind(a, "H5Eset_auto2", &L.eset_auto2) &&
            bind(b, "H5LTmake_dataset_double", &L.make_double) && bind(b, "H5LTread_dataset_double", &L.read_double) &&
            bind(b, "H5LTset_attribute_int", &L.set_attr_int) && bind(b, "H5LTget_attribute_int", &L.get_attr_int) &&
            bind(b, "H5LTfind_dataset", &L.find_dataset) && bind(b, "H5LTget_dataset_ndims", &L.get_ndims) &&
            bind(b, "H5LTget_dataset_info", &L.get_info);
  if (!ok || L.open() < 0) { L.err = "HDF5 library lacks an expected symbol"; return L; }
  L.eset_auto2(0, nullptr, nullptr);   // no console error messages, as the reference
  L.ok = true;
  return L;
}

struct Ctx {
  hid_t file;
};

// opens a file for reading and checks the dataset's rank
std::string open_dataset(const char *path, const char *name, int rank, hid_t *file, hsize_t dims[2])
{
  H5 &L = h5();
  if (!L.ok) return L.err;
  *file = L.fopen(path, kAccRdonly, kDefault);
  if (*file < 0) return std::string("could not open file ") + path;
  int r = 0, cls = 0;
  size_t sz = 0;
  if (L.find_dataset(*file, name) <= 0) { L.fclose(*file); return std::string("file ") + path + " does not contain a dataset named " + name; }
  if (L.get_ndims(*file, name, &r) < 0 || r != rank) {
    L.fclose(*file);
    return std::string("file ") + path + ": dataset " + name + (rank == 2 ? " is not a matrix" : " is not a vector");
  }
  if (L.get_info(*file, name, dims, &cls, &sz) < 0) { L.fclose(*file); return std::string("file ") + path + ": dataset " + name + " has incorrect type"; }
  return "";
}

}  // namespace

std::string hdf5_open(const char *path, void **ctx)
{
  H5 &L = h5();
  if (!L.ok) return L.err;
  hid_t f = L.fcreate(path, kAccTrunc, kDefault, kDefault);
  if (f < 0) return std::string("could not open file ") + path;
  *ctx = new Ctx{f};
  return "";
}

std::string hdf5_close(void *ctx)
{
  H5 &L = h5();
  Ctx *c = (Ctx *)ctx;
  if (!L.ok || !c) return L.ok ? "null context" : L.err;
  L.fclose(c->file);
  delete c;
  return "";
}

std::string hdf5_write_matrix(void *ctx, const char *name, const double *data, int nr, int nc, int cplx)
{
  H5 &L = h5();
  Ctx *c = (Ctx *)ctx;
  if (!L.ok) return L.err;
  hsize_t dims[2] = {(hsize_t)nc, (hsize_t)(cplx ? 2 * nr : nr)};
  if (L.make_double(c->file, name, 2, dims, data) != 0) return std::string("could not write dataset ") + name;
  int st = 0, rc = cplx ? 1 : 0;
  L.set_attr_int(c->file, name, "Storage_Type", &st, 1);
  L.set_attr_int(c->file, name, "RealComplex", &rc, 1);
  return "";
}

std::string hdf5_write_vector(void *ctx, const char *name, const double *data, int n, int cplx)
{
  H5 &L = h5();
  Ctx *c = (Ctx *)ctx;
  if (!L.ok) return L.err;
  hsize_t dims[1] = {(hsize_t)(cplx ? 2 * n : n)};
  if (L.make_double(c->file, name, 1, dims, data) != 0) return std::string("could not write dataset ") + name;
  int rc = cplx ? 1 : 0;
  L.set_attr_int(c->file, name, "RealComplex", &rc, 1);
  return "";
}

std::string hdf5_read_matrix(const char *path, const char *name, double *data, int *nr, int *nc, int *cplx)
{
  H5 &L = h5();
  hid_t f;
  hsize_t dims[2] = {0, 0};
  std::string e = open_dataset(path, name, 2, &f, dims);
  if (!e.empty()) return e;
  // without the attributes (a file from another application) the data
  // is taken to be a real matrix in normal storage
  int st = 0, rc = 0;
  L.get_attr_int(f, name, "Storage_Type", &st);
  L.get_attr_int(f, name, "RealComplex", &rc);
  if (st != 0) { L.fclose(f); return std::string("file ") + path + ", dataset " + name + ": packed storage is not supported"; }
  *cplx = rc ? 1 : 0;
  *nc = (int)dims[0];
  *nr = (int)(rc ? dims[1] / 2 : dims[1]);
  if (data && L.read_double(f, name, data) < 0) { L.fclose(f); return std::string("could not read dataset ") + name; }
  L.fclose(f);
  return "";
}

std::string hdf5_read_vector(const char *path, const char *name, double *data, int *n, int *cplx)
{
  H5 &L = h5();
  hid_t f;
  hsize_t dims[2] = {0, 0};
  std::string e = open_dataset(path, name, 1, &f, dims);
  if (!e.empty()) return e;
  int rc = 0;
  if (L.get_attr_int(f, name, "RealComplex", &rc) < 0) rc = 0;
  *cplx = rc ? 1 : 0;
  *n = (int)(rc ? dims[0] / 2 : dims[0]);
  if (data && L.read_double(f, name, data) < 0) { L.fclose(f); return std::string("could not read dataset ") + name; }
  L.fclose(f);
  return "";
}

}  // namespace scuffx
