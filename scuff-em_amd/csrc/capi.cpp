// capi.cpp -- extern "C" boundary (include/scuffx.h) over the host
// geometry, the assembly orchestrator and the LU module.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/scuffx.h"
#include "assemble.h"
#include "fippi_cache.h"
#include "geometry.h"
#include "hdf5_io.h"
#include "adjoint.h"
#include "lu.h"
#include "prof.h"
#include "refine.h"
#include "rhs.h"
#include "ws.h"

using namespace scuffx;

static thread_local std::string g_err;
static thread_local hipStream_t g_stream = nullptr;

static int fail(int code, const std::string &msg) { g_err = msg; return code; }
static int refine_status(int rc, double rel, double tol);

// for entry points defined in other files of the library
namespace scuffx {
int capi_fail(int code, const char *msg) { return fail(code, msg); }
int capi_refine_status(int rc, double rel, double tol) { return refine_status(rc, rel, tol); }
hipStream_t capi_stream() { return g_stream; }
}  // namespace scuffx

struct DevRhs {
  double *PA = nullptr, *PZ = nullptr, *work = nullptr;
  int *PEI = nullptr, *EV = nullptr;
};

struct scuffx_trans {
  std::vector<GTComplex> C;
};

struct scuffx_geom;
namespace {
void sweep_release_all();   // the batched sweep's kept buffers (below)
}
// live geometries, so that scuffx_shutdown can release their device memory
static std::mutex g_geom_mu;
static std::set<scuffx_geom *> &live_geoms()
{
  static std::set<scuffx_geom *> *s = new std::set<scuffx_geom *>();   // never destroyed
  return *s;
}

struct scuffx_geom {
  Geometry G;
  std::unique_ptr<Assembler> A;
  bool frozen = false;          // surfaces fixed once an assembler existed
  bool transformed = false;     // a surface moved since the last UnTransform
  int far_bits = 64;            // precision of the far panel pairs (32 or 64)
  std::vector<DevRhs> rhs;
  double *d_out = nullptr;
  signed char *d_bft = nullptr;   // basis-function types for the EMT power
  double *d_emt = nullptr;        // its column partial sums (TotalBFs + 1)
  scuffx_geom()
  {
    std::lock_guard<std::mutex> lk(g_geom_mu);
    live_geoms().insert(this);
  }
  ~scuffx_geom()
  {
    {
      std::lock_guard<std::mutex> lk(g_geom_mu);
      live_geoms().erase(this);
    }
    release_device();
  }
  // frees every device buffer (no HIP call after shutdown)
  void release_device()
  {
    if (hip_down()) return;
    A.reset();
    drop_rhs();
    hipFree(d_out);
    d_out = nullptr;
  }
  void drop_rhs()
  {
    if (!hip_down()) {
      for (auto &r : rhs) { hipFree(r.PA); hipFree(r.PZ); hipFree(r.work); hipFree(r.PEI); hipFree(r.EV); }
      hipFree(d_bft);
      hipFree(d_emt);
    }
    d_bft = nullptr;
    d_emt = nullptr;
    rhs.clear();
  }
  // the geometry moved: device copies are rebuilt on their next use
  void moved()
  {
    if (A || !rhs.empty()) hipStreamSynchronize(g_stream);
    A.reset();
    drop_rhs();
  }
  Assembler *assembler()
  {
    if (!A) {
      A.reset(new Assembler(G));
      A->set_far_precision(far_bits);
    }
    frozen = true;
    return A.get();
  }
};

extern "C" {

const char *scuffx_last_error(void) { return g_err.c_str(); }
int scuffx_version(void) { return 200; }

int scuffx_shutdown(void)
{
  if (hip_down()) return 0;
  hipDeviceSynchronize();
  {
    std::lock_guard<std::mutex> lk(g_geom_mu);
    for (scuffx_geom *g : live_geoms()) g->release_device();
  }
  sweep_release_all();
  ws_release_all();
  set_hip_down();
  return 0;
}
int scuffx_set_stream(void *s) { g_stream = (hipStream_t)s; return 0; }

int scuffx_geom_new(scuffx_geom **G)
{
  if (!G) return fail(SCUFFX_ERR_ARG, "null handle pointer");
  *G = new scuffx_geom();
  return 0;
}

int scuffx_geom_open(const char *path, scuffx_geom **G)
{
  if (!path || !G) return fail(SCUFFX_ERR_ARG, "null argument");
  scuffx_geom *g = new scuffx_geom();
  std::string e = read_scuffgeo(path, &g->G);
  if (!e.empty()) { delete g; return fail(SCUFFX_ERR_IO, e); }
  *G = g;
  return 0;
}

static int add_surface_checked(scuffx_geom *G, const double *V, int nV, const int *T, int nT, const char *mat,
                               const char *meshfile = nullptr)
{
  if (G->frozen) return fail(SCUFFX_ERR_ARG, "geometry is frozen after its first assembly");
  Material m;
  bool pec;
  std::string e = parse_material(mat ? mat : "PEC", &m, &pec);
  if (!e.empty()) return fail(SCUFFX_ERR_ARG, e);
  Surface S;
  e = build_surface(V, nV, T, nT, pec, &S);
  if (!e.empty()) return fail(SCUFFX_ERR_ARG, e);
  S.MeshFile = meshfile ? meshfile : "";
  S.MaterialName = pec ? std::string("PEC") : std::string(mat);
  S.Label = meshfile ? meshfile : "ByHand";
  return add_surface(&G->G, std::move(S), m);
}

int scuffx_geom_add_mesh(scuffx_geom *G, const double *V, int nV, const int *T, int nT, const char *mat)
{
  if (!G || !V || !T || nV <= 0 || nT <= 0) return fail(SCUFFX_ERR_ARG, "bad mesh arguments");
  return add_surface_checked(G, V, nV, T, nT, mat);
}

int scuffx_geom_add_msh(scuffx_geom *G, const char *path, const char *mat)
{
  if (!G || !path) return fail(SCUFFX_ERR_ARG, "null argument");
  std::vector<double> V;
  std::vector<int> T;
  std::string e = read_gmsh(path, &V, &T);
  if (!e.empty()) return fail(SCUFFX_ERR_IO, e);
  return add_surface_checked(G, V.data(), (int)V.size() / 3, T.data(), (int)T.size() / 3, mat, path);
}

int scuffx_geom_mate(const scuffx_geom *G, int ns)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  return find_mates(G->G)[ns];
}

int scuffx_geom_epsmu(const scuffx_geom *G, int region, double wr, double wi, double *eps2, double *mu2)
{
  if (!G || region < 0 || region >= (int)G->G.Regions.size()) return fail(SCUFFX_ERR_ARG, "bad region index");
  cdouble e, m;
  G->G.Regions[region].eps_mu(cdouble(wr, wi), &e, &m);
  if (eps2) { eps2[0] = e.real(); eps2[1] = e.imag(); }
  if (mu2) { mu2[0] = m.real(); mu2[1] = m.imag(); }
  return 0;
}

int scuffx_geom_free(scuffx_geom *G) { delete G; return 0; }
int scuffx_geom_total_bfs(const scuffx_geom *G) { return G ? G->G.TotalBFs : fail(SCUFFX_ERR_ARG, "null geometry"); }
int scuffx_geom_num_surfaces(const scuffx_geom *G) { return G ? (int)G->G.Surfaces.size() : fail(SCUFFX_ERR_ARG, "null geometry"); }

int scuffx_geom_counts(const scuffx_geom *G, int ns, int *nV, int *nP, int *nE)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  const Surface &S = G->G.Surfaces[ns];
  if (nV) *nV = S.NumVertices();
  if (nP) *nP = S.NumPanels();
  if (nE) *nE = S.NumEdges();
  return 0;
}

int scuffx_geom_bf_range(const scuffx_geom *G, int ns, int *offset, int *nbf)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  if (offset) *offset = G->G.BFIndexOffset[ns];
  if (nbf) *nbf = G->G.Surfaces[ns].NumBFs;
  return 0;
}

int scuffx_geom_edges(const scuffx_geom *G, int ns, int *out, double *len)
{
  if (!G || !out || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  const Surface &S = G->G.Surfaces[ns];
  for (int e = 0; e < S.NumEdges(); e++) {
    const Edge &E = S.Edges[e];
    int *o = out + 9 * e;
    o[0] = E.iV1; o[1] = E.iV2; o[2] = E.iQP; o[3] = E.iQM; o[4] = E.iPPanel; o[5] = E.iMPanel;
    o[6] = E.PIndex; o[7] = E.MIndex; o[8] = E.Index;
    if (len) len[e] = E.Length;
  }
  return 0;
}

int scuffx_geom_panels(const scuffx_geom *G, int ns, double *out)
{
  if (!G || !out || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  const Surface &S = G->G.Surfaces[ns];
  for (int p = 0; p < S.NumPanels(); p++) {
    const Panel &P = S.Panels[p];
    double *o = out + 8 * p;
    for (int j = 0; j < 3; j++) { o[j] = P.Centroid[j]; o[3 + j] = P.ZHat[j]; }
    o[6] = P.Radius; o[7] = P.Area;
  }
  return 0;
}

int scuffx_geom_vertices(const scuffx_geom *G, int ns, double *out)
{
  if (!G || !out || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  const Surface &S = G->G.Surfaces[ns];
  for (size_t i = 0; i < S.V.size(); i++) out[i] = S.V[i];
  return 0;
}

static int copy_str(const std::string &v, char *buf, int len)
{
  if (!buf || len <= 0) return fail(SCUFFX_ERR_ARG, "no buffer");
  std::snprintf(buf, (size_t)len, "%s", v.c_str());
  return (int)v.size() < len ? 0 : fail(SCUFFX_ERR_ARG, "buffer too short");
}

int scuffx_trans_open(const char *path, scuffx_trans **T)
{
  if (!T) return fail(SCUFFX_ERR_ARG, "null handle pointer");
  scuffx_trans *t = new scuffx_trans();
  std::string e = read_trans_file(path ? path : "", &t->C);
  if (!e.empty()) { delete t; return fail(SCUFFX_ERR_IO, e); }
  *T = t;
  return 0;
}

int scuffx_trans_count(const scuffx_trans *T) { return T ? (int)T->C.size() : fail(SCUFFX_ERR_ARG, "null handle"); }

int scuffx_trans_tag(const scuffx_trans *T, int i, char *buf, int len)
{
  if (!T || i < 0 || i >= (int)T->C.size()) return fail(SCUFFX_ERR_ARG, "bad transformation index");
  return copy_str(T->C[i].Tag, buf, len);
}

int scuffx_trans_free(scuffx_trans *T) { delete T; return 0; }

int scuffx_geom_label(const scuffx_geom *G, int ns, char *buf, int len)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  return copy_str(G->G.Surfaces[ns].Label, buf, len);
}

int scuffx_geom_set_label(scuffx_geom *G, int ns, const char *label)
{
  if (!G || !label || !*label || ns < 0 || ns >= (int)G->G.Surfaces.size())
    return fail(SCUFFX_ERR_ARG, "bad surface index or label");
  G->G.Surfaces[ns].Label = label;
  return 0;
}

int scuffx_geom_check_trans(const scuffx_geom *G, const scuffx_trans *T)
{
  if (!G || !T) return fail(SCUFFX_ERR_ARG, "null argument");
  for (const GTComplex &C : T->C) {
    std::string e = check_gtc(G->G, C);
    if (!e.empty()) return fail(SCUFFX_ERR_ARG, e);
  }
  return 0;
}

int scuffx_geom_transform(scuffx_geom *G, const scuffx_trans *T, int i)
{
  if (!G || !T || i < 0 || i >= (int)T->C.size()) return fail(SCUFFX_ERR_ARG, "bad transformation index");
  std::string e = check_gtc(G->G, T->C[i]);
  if (!e.empty()) return fail(SCUFFX_ERR_ARG, e);
  transform_geometry(&G->G, T->C[i]);
  for (int m : G->G.SurfaceMoved)
    if (m) { G->moved(); G->transformed = true; break; }
  return 0;
}

int scuffx_geom_transform_surface(scuffx_geom *G, int ns, const double *R9, const double *D3)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  GTransform T;
  for (int i = 0; i < 3; i++) {
    if (R9) for (int j = 0; j < 3; j++) T.M[i][j] = R9[3 * i + j];
    if (D3) T.DX[i] = D3[i];
  }
  G->G.SurfaceMoved.assign(G->G.Surfaces.size(), 0);
  transform_surface(&G->G, ns, T);
  G->moved();
  G->transformed = true;
  return 0;
}

int scuffx_geom_set_far_precision(scuffx_geom *G, int bits)
{
  if (!G || (bits != 32 && bits != 64)) return fail(SCUFFX_ERR_ARG, "precision must be 32 or 64");
  G->far_bits = bits;
  if (G->A) G->A->set_far_precision(bits);
  return 0;
}

int scuffx_geom_set_vertices(scuffx_geom *G, int ns, const double *V, int nV)
{
  if (!G || !V || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (G->transformed) return fail(SCUFFX_ERR_ARG, "the geometry is transformed: UnTransform first");
  const Surface &S = G->G.Surfaces[ns];
  if (nV != S.NumVertices()) return fail(SCUFFX_ERR_ARG, "vertex count differs from the mesh");
  if (std::equal(V, V + 3 * (size_t)nV, S.V.begin())) return 0;   // unchanged: keep the device state
  if (G->A || !G->rhs.empty()) hipStreamSynchronize(g_stream);
  std::string e = set_surface_vertices(&G->G, ns, V, nV);
  if (!e.empty()) return fail(SCUFFX_ERR_ARG, e);
  G->moved();
  return 0;
}

int scuffx_geom_untransform(scuffx_geom *G)
{
  if (!G) return fail(SCUFFX_ERR_ARG, "null geometry");
  if (!G->transformed) return 0;    // nothing moved since the last UnTransform
  untransform_geometry(&G->G);
  G->moved();
  G->transformed = false;
  return 0;
}

int scuffx_geom_surface_moved(const scuffx_geom *G, int ns)
{
  if (!G || ns < 0 || ns >= (int)G->G.Surfaces.size()) return fail(SCUFFX_ERR_ARG, "bad surface index");
  return ns < (int)G->G.SurfaceMoved.size() ? G->G.SurfaceMoved[ns] : 0;
}

static int io_status(const std::string &e) { return e.empty() ? 0 : fail(SCUFFX_ERR_IO, e); }

int scuffx_hdf5_open(const char *path, void **ctx)
{
  if (!path || !ctx) return fail(SCUFFX_ERR_ARG, "null argument");
  return io_status(hdf5_open(path, ctx));
}

int scuffx_hdf5_close(void *ctx)
{
  if (!ctx) return fail(SCUFFX_ERR_ARG, "null context");
  return io_status(hdf5_close(ctx));
}

int scuffx_hdf5_write_matrix(void *ctx, const char *name, const double *data, int nr, int nc, int is_complex)
{
  if (!ctx || !name || !data || nr < 0 || nc < 0) return fail(SCUFFX_ERR_ARG, "bad arguments");
  return io_status(hdf5_write_matrix(ctx, name, data, nr, nc, is_complex ? 1 : 0));
}

int scuffx_hdf5_write_vector(void *ctx, const char *name, const double *data, int n, int is_complex)
{
  if (!ctx || !name || !data || n < 0) return fail(SCUFFX_ERR_ARG, "bad arguments");
  return io_status(hdf5_write_vector(ctx, name, data, n, is_complex ? 1 : 0));
}

int scuffx_hdf5_read_matrix(const char *path, const char *name, double *data, int *nr, int *nc, int *is_complex)
{
  if (!path || !name || !nr || !nc || !is_complex) return fail(SCUFFX_ERR_ARG, "bad arguments");
  return io_status(hdf5_read_matrix(path, name, data, nr, nc, is_complex));
}

int scuffx_hdf5_read_vector(const char *path, const char *name, double *data, int *n, int *is_complex)
{
  if (!path || !name || !n || !is_complex) return fail(SCUFFX_ERR_ARG, "bad arguments");
  return io_status(hdf5_read_vector(path, name, data, n, is_complex));
}

int scuffx_icosphere(int level, double radius, const double *center, double *V, int *T)
{
  if (level < 0 || level > 9 || !V || !T) return fail(SCUFFX_ERR_ARG, "bad icosphere arguments");
  std::vector<double> v;
  std::vector<int> t;
  icosphere(level, radius, center, &v, &t);
  std::memcpy(V, v.data(), sizeof(double) * v.size());
  std::memcpy(T, t.data(), sizeof(int) * t.size());
  return 0;
}

int scuffx_fippi_build(scuffx_geom *G)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || G->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "empty geometry");
  Assembler *A = G->assembler();
  int rc = A->build_fippi(g_stream);
  if (rc) return fail(rc, A->error());
  return 0;
}

int scuffx_fippi_count(const scuffx_geom *G)
{
  if (!G) return fail(SCUFFX_ERR_ARG, "null geometry");
  return G->A ? G->A->num_fippi() : 0;
}

long scuffx_preload_cache(const char *path)
{
  if (!path) return fail(SCUFFX_ERR_ARG, "null file name");
  std::string err;
  long n = fippi_cache_preload(path, err);
  if (n < 0) return fail(SCUFFX_ERR_ARG, err + " (cache not preloaded)");
  return n;
}

long scuffx_store_cache(const char *path)
{
  if (!path) return fail(SCUFFX_ERR_ARG, "null file name");
  std::string err;
  long n = fippi_cache_store(path, err);
  if (n < 0) return fail(SCUFFX_ERR_ARG, err);
  return n;
}

long scuffx_cache_size(void) { return (long)fippi_cache_size(); }

int scuffx_clear_cache(void)
{
  fippi_cache_clear();
  return 0;
}

int scuffx_fippi_records(scuffx_geom *G, double *out)
{
  if (!G || !out) return fail(SCUFFX_ERR_ARG, "null argument");
  if (!G->A) return 0;
  std::vector<double> r = G->A->fippi_records();
  std::memcpy(out, r.data(), sizeof(double) * r.size());
  return 0;
}

int scuffx_assemble(scuffx_geom *G, double wr, double wi, void *M, long ldm)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !M || G->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  Assembler *A = G->assembler();
  int rc = A->assemble(cdouble(wr, wi), M, ldm, MAT_C128, g_stream);
  if (rc) return fail(rc, A->error());
  return 0;
}

int scuffx_assemble_real(scuffx_geom *G, double wr, double wi, double *M, long ldm, int packed)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !M || G->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  Assembler *A = G->assembler();
  int rc = A->assemble(cdouble(wr, wi), M, ldm, packed ? MAT_F64_PACKED : MAT_F64, g_stream);
  if (rc) return fail(rc, A->error());
  return 0;
}

int scuffx_assemble_block_real(scuffx_geom *G, int nsa, int nsb, double wr, double wi, double *B, long ldb,
                               int packed, double *GradB, long ldg, long gstride, int ntorque, const double *gamma,
                               double *dBdT)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !B || G->G.Surfaces.empty() || nsa < 0 || nsb < 0 || nsa >= (int)G->G.Surfaces.size() ||
      nsb >= (int)G->G.Surfaces.size() || ntorque < 0 || ntorque > 3)
    return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (packed && (nsa != nsb || GradB)) return fail(SCUFFX_ERR_ARG, "packed storage is for diagonal blocks only");
  const long na = G->G.Surfaces[nsa].NumBFs, nb = G->G.Surfaces[nsb].NumBFs;
  if ((!packed && ldb < na) || (GradB && (ldg < na || gstride < ldg * nb)))
    return fail(SCUFFX_ERR_ARG, "leading dimension smaller than the block");
  // the complex block (and its derivatives) are staged, then their real
  // parts stored: at an imaginary frequency the imaginary parts vanish
  const long nmat = 1 + (GradB ? 3 + ntorque : 0);
  double2 *st = (double2 *)ws_get(WS_STAGE, sizeof(double2) * na * nb * nmat);
  if (!st) return fail(SCUFFX_ERR_ALLOC, "staging allocation failed");
  const long bs = na * nb;
  int rc = scuffx_assemble_block(G, nsa, nsb, wr, wi, st, na, GradB ? st + bs : nullptr, na, bs, ntorque, gamma,
                                 ntorque ? st + 4 * bs : nullptr);
  if (rc) return rc;
  if (real_part_c128(st, na, B, ldb, na, nb, packed, g_stream)) return fail(SCUFFX_ERR_HIP, "conversion failed");
  for (long m = 1; m < nmat; m++) {
    double *dst = m <= 3 ? GradB + (m - 1) * gstride : dBdT + (m - 4) * gstride;
    if (real_part_c128(st + m * bs, na, dst, ldg, na, nb, 0, g_stream)) return fail(SCUFFX_ERR_HIP, "conversion failed");
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "conversion failed");
  return 0;
}

int scuffx_unpack_symmetric(const void *P, long n, void *A, long lda, int prec)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!P || !A || n < 0 || lda < (n > 1 ? n : 1)) return fail(SCUFFX_ERR_ARG, "bad arguments");
  int rc;
  if (prec == 8) rc = unpack_symmetric<double>((const double *)P, (double *)A, lda, n, g_stream);
  else if (prec == 64) rc = unpack_symmetric<double2>((const double2 *)P, (double2 *)A, lda, n, g_stream);
  else if (prec == 32) rc = unpack_symmetric<float2>((const float2 *)P, (float2 *)A, lda, n, g_stream);
  else return fail(SCUFFX_ERR_ARG, "prec must be 64, 32 or 8");
  if (rc || hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "unpack failed");
  return 0;
}

int scuffx_assemble_block(scuffx_geom *G, int nsa, int nsb, double wr, double wi, void *B, long ldb, void *GradB,
                          long ldg, long gstride, int ntorque, const double *gamma, void *dBdT)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !B || G->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  Assembler *A = G->assembler();
  int rc = A->assemble_block(nsa, nsb, cdouble(wr, wi), (double2 *)B, ldb, (double2 *)GradB, ldg, gstride, ntorque,
                             gamma, (double2 *)dBdT, g_stream);
  if (rc) return fail(rc, A->error());
  return 0;
}

int scuffx_assemble_stats(const scuffx_geom *G, double *ms, long *counts)
{
  if (!G || !G->A) return fail(SCUFFX_ERR_ARG, "no assembly yet");
  const AsmStats &s = G->A->stats();
  if (ms) { ms[0] = s.fippi_ms; ms[1] = s.self_ms; ms[2] = s.far_ms; ms[3] = s.near_ms; ms[4] = s.stamp_ms; ms[5] = s.mirror_ms; }
  if (counts) { counts[0] = s.nfar; counts[1] = s.nnear; counts[2] = s.nself; counts[3] = s.nfippi; }
  return 0;
}

static int lu_status(int rc, int info)
{
  if (rc == -5) return fail(SCUFFX_ERR_UNSUPPORTED, "matrix too large for a co-resident panel grid");
  if (rc == -6) return fail(SCUFFX_ERR_HIP, "panel grid barrier timed out");
  if (rc == -2) return fail(SCUFFX_ERR_ALLOC, "LU workspace allocation failed");
  if (rc) return fail(SCUFFX_ERR_HIP, "LU factorization failed on the device");
  return info;
}

int scuffx_lu_factor(void *A, long n, long lda, int *ipiv, int prec)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!A || !ipiv || n < 0 || lda < (n > 1 ? n : 1)) return fail(SCUFFX_ERR_ARG, "bad LU arguments");
  int info = 0, rc;
  if (prec == 64) rc = lu_factor<double2>((double2 *)A, n, lda, ipiv, &info, g_stream);
  else if (prec == 32) rc = lu_factor<float2>((float2 *)A, n, lda, ipiv, &info, g_stream);
  else if (prec == 8) rc = lu_factor<double>((double *)A, n, lda, ipiv, &info, g_stream);
  else return fail(SCUFFX_ERR_ARG, "prec must be 64, 32 or 8");
  return lu_status(rc, info);
}

int scuffx_lu_solve(const void *LU, long n, long lda, const int *ipiv, char trans, long nrhs, void *B, long ldb, int prec)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!LU || !ipiv || !B) return fail(SCUFFX_ERR_ARG, "null argument");
  int rc;
  if (prec == 64) rc = lu_solve<double2>((const double2 *)LU, n, lda, ipiv, trans, nrhs, (double2 *)B, ldb, g_stream);
  else if (prec == 32) rc = lu_solve<float2>((const float2 *)LU, n, lda, ipiv, trans, nrhs, (float2 *)B, ldb, g_stream);
  else if (prec == 8) rc = lu_solve<double>((const double *)LU, n, lda, ipiv, trans, nrhs, (double *)B, ldb, g_stream);
  else return fail(SCUFFX_ERR_ARG, "prec must be 64, 32 or 8");
  if (rc == -1) return fail(SCUFFX_ERR_ARG, "bad LU solve arguments");
  if (rc == -2) return fail(SCUFFX_ERR_ALLOC, "solve workspace allocation failed");
  if (rc) return fail(SCUFFX_ERR_HIP, "LU solve failed on the device");
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "LU solve failed on the device");
  return 0;
}

int scuffx_lu_solve_noalloc(const void *LU, long n, long lda, const int *ipiv, void *B, int prec)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!LU || !ipiv || !B) return fail(SCUFFX_ERR_ARG, "null argument");
  int rc;
  if (prec == 64) rc = lu_solve_noalloc<double2>((const double2 *)LU, n, lda, ipiv, (double2 *)B, g_stream);
  else if (prec == 32) rc = lu_solve_noalloc<float2>((const float2 *)LU, n, lda, ipiv, (float2 *)B, g_stream);
  else return fail(SCUFFX_ERR_ARG, "prec must be 64 or 32");
  if (rc == -7) return fail(SCUFFX_ERR_ARG, "n exceeds the no-allocation solve's maximum of 16384");
  if (rc == -1) return fail(SCUFFX_ERR_ARG, "bad LU solve arguments");
  if (rc) return fail(SCUFFX_ERR_HIP, "LU solve failed on the device");
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "LU solve failed on the device");
  return 0;
}

long scuffx_lu_solve_noalloc_bytes(void) { return noalloc_scratch_bytes(); }

int scuffx_lu_logdet(const void *LU, long n, long lda, const int *ipiv, int prec, double *logabs, double *phase)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!LU || !ipiv || n < 0 || lda < (n > 1 ? n : 1)) return fail(SCUFFX_ERR_ARG, "bad arguments");
  double out[2] = {0, 0};
  int rc;
  if (prec == 64) rc = lu_logdet<double2>((const double2 *)LU, n, lda, ipiv, out, g_stream);
  else if (prec == 32) rc = lu_logdet<float2>((const float2 *)LU, n, lda, ipiv, out, g_stream);
  else if (prec == 8) rc = lu_logdet<double>((const double *)LU, n, lda, ipiv, out, g_stream);
  else return fail(SCUFFX_ERR_ARG, "prec must be 64, 32 or 8");
  if (rc) return fail(SCUFFX_ERR_HIP, "log det failed on the device");
  if (logabs) *logabs = out[0];
  if (phase) *phase = out[1];
  return 0;
}

int scuffx_trace_solve(const void *LU, long n, long lda, const int *ipiv, void *D, long ldd, long ncols, int prec,
                       double *tr_re, double *tr_im)
{
  if (!LU || !ipiv || !D || n <= 0 || ncols <= 0 || ncols > n) return fail(SCUFFX_ERR_ARG, "bad arguments");
  int rc = scuffx_lu_solve(LU, n, lda, ipiv, 'N', ncols, D, ldd, prec);
  if (rc) return rc;
  double out[2] = {0, 0};
  if (prec == 64) rc = mat_trace<double2>((const double2 *)D, ncols, ldd, out, g_stream);
  else if (prec == 8) rc = mat_trace<double>((const double *)D, ncols, ldd, out, g_stream);
  else rc = mat_trace<float2>((const float2 *)D, ncols, ldd, out, g_stream);
  if (rc) return fail(SCUFFX_ERR_HIP, "trace failed on the device");
  if (tr_re) *tr_re = out[0];
  if (tr_im) *tr_im = out[1];
  return 0;
}

int scuffx_lu_invert(void *A, long n, long lda, const int *ipiv, int prec)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!A || !ipiv || n < 0 || lda < (n > 1 ? n : 1)) return fail(SCUFFX_ERR_ARG, "bad arguments");
  size_t es = prec == 64 ? sizeof(double2) : prec == 32 ? sizeof(float2) : prec == 8 ? sizeof(double) : 0;
  if (!es) return fail(SCUFFX_ERR_ARG, "prec must be 64, 32 or 8");
  void *W = ws_get(WS_STAGE, es * (size_t)n * n + 256);
  if (!W) return fail(SCUFFX_ERR_ALLOC, "workspace allocation failed");
  int info = 0, rc;
  if (prec == 64) rc = lu_invert<double2>((double2 *)A, n, lda, ipiv, (double2 *)W, &info, g_stream);
  else if (prec == 32) rc = lu_invert<float2>((float2 *)A, n, lda, ipiv, (float2 *)W, &info, g_stream);
  else rc = lu_invert<double>((double *)A, n, lda, ipiv, (double *)W, &info, g_stream);
  if (rc == -2) return fail(SCUFFX_ERR_ALLOC, "solve workspace allocation failed");
  if (rc) return fail(SCUFFX_ERR_HIP, "LU inverse failed on the device");
  return info;
}

int scuffx_lu32_from_c128(const void *M, long n, long ldm, void *LU32, long ld32, int *ipiv)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!M || !LU32 || !ipiv || ldm < n || ld32 < n) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (convert_c128_to_c64((const double2 *)M, (float2 *)LU32, n, n, ldm, ld32, g_stream)) return fail(SCUFFX_ERR_HIP, "conversion failed");
  int info = 0;
  int rc = lu_factor<float2>((float2 *)LU32, n, ld32, ipiv, &info, g_stream, 256, 1);
  return lu_status(rc, info);
}

static int refine_status(int rc, double rel, double tol)
{
  if (rc == 1) {
    char b[160];
    std::snprintf(b, sizeof b, "refinement stopped at relative residual %.3e above the tolerance %.3e", rel, tol);
    g_err = b;
    return SCUFFX_NOT_CONVERGED;
  }
  if (rc == -1) return fail(SCUFFX_ERR_ARG, "bad refinement arguments");
  if (rc == -2) return fail(SCUFFX_ERR_ALLOC, "refinement workspace allocation failed");
  if (rc == -4) return fail(SCUFFX_ERR_NUMERIC, "non-finite residual in the refinement");
  if (rc == -6) return fail(SCUFFX_ERR_HIP, "a triangular solve's bounded wait timed out");
  if (rc) return fail(SCUFFX_ERR_HIP, "refinement failed on the device");
  return 0;
}

static int refine_common(const void *M, long n, long ldm, const void *LU32, long ld32, const int *ipiv,
                         const void *b, void *x, double tol, int maxiter, int *iters, double *relres, int mode)
{
  if (!M || !LU32 || !ipiv || !b || !x || n <= 0 || ldm < n || ld32 < n || maxiter < 0)
    return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  char *wsb = (char *)ws_get(WS_REFINE, refine_ws_bytes(n, REFINE_GMRES_M));
  if (!wsb) return fail(SCUFFX_ERR_ALLOC, "workspace allocation failed");
  RefineWs w = refine_ws_carve(wsb, n, REFINE_GMRES_M);
  int it = 0;
  double rel = 0;
  int rc = mode == 0 ? refine_ir((const double2 *)M, n, ldm, (const float2 *)LU32, ld32, ipiv, (const double2 *)b,
                                 (double2 *)x, tol, maxiter, 1, &it, &rel, w, g_stream)
                     : gmres_ir((const double2 *)M, n, ldm, (const float2 *)LU32, ld32, ipiv, (const double2 *)b,
                                (double2 *)x, tol, maxiter, false, &it, &rel, w, g_stream);
  if (prof_enabled()) prof_collect();
  if (iters) *iters = it;
  if (relres) *relres = rel;
  return refine_status(rc, rel, tol);
}

int scuffx_refine(const void *M, long n, long ldm, const void *LU32, long ld32, const int *ipiv,
                  const void *b, void *x, double tol, int maxiter, int *iters, double *relres)
{
  return refine_common(M, n, ldm, LU32, ld32, ipiv, b, x, tol, maxiter, iters, relres, 0);
}

int scuffx_gmres_ir(const void *M, long n, long ldm, const void *LU32, long ld32, const int *ipiv,
                    const void *b, void *x, double tol, int maxiter, int *iters, double *relres)
{
  return refine_common(M, n, ldm, LU32, ld32, ipiv, b, x, tol, maxiter, iters, relres, 1);
}

static int ensure_rhs_data(scuffx_geom *G)
{
  if (!G->rhs.empty()) return 0;
  for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
    const Surface &S = G->G.Surfaces[s];
    DevRhs R;
    std::vector<double> pa(S.NumPanels()), pz(3 * S.NumPanels());
    std::vector<int> pei(3 * S.NumPanels()), ev(4 * S.NumEdges());
    for (int p = 0; p < S.NumPanels(); p++) {
      pa[p] = S.Panels[p].Area;
      for (int j = 0; j < 3; j++) { pz[3 * p + j] = S.Panels[p].ZHat[j]; pei[3 * p + j] = S.Panels[p].EI[j]; }
    }
    for (int e = 0; e < S.NumEdges(); e++) {
      ev[4 * e] = S.Edges[e].iV1; ev[4 * e + 1] = S.Edges[e].iV2; ev[4 * e + 2] = S.Edges[e].iQP; ev[4 * e + 3] = S.Edges[e].iQM;
    }
    if (hipMalloc(&R.PA, sizeof(double) * pa.size()) || hipMalloc(&R.PZ, sizeof(double) * pz.size()) ||
        hipMalloc(&R.PEI, sizeof(int) * pei.size()) || hipMalloc(&R.EV, sizeof(int) * std::max<size_t>(1, ev.size())) ||
        hipMalloc(&R.work, sizeof(double) * 2 * std::max(1, S.NumEdges())))
      return fail(SCUFFX_ERR_ALLOC, "allocation failed");
    hipMemcpy(R.PA, pa.data(), sizeof(double) * pa.size(), hipMemcpyHostToDevice);
    hipMemcpy(R.PZ, pz.data(), sizeof(double) * pz.size(), hipMemcpyHostToDevice);
    hipMemcpy(R.PEI, pei.data(), sizeof(int) * pei.size(), hipMemcpyHostToDevice);
    if (!ev.empty()) hipMemcpy(R.EV, ev.data(), sizeof(int) * ev.size(), hipMemcpyHostToDevice);
    G->rhs.push_back(R);
  }
  if (!G->d_out && hipMalloc(&G->d_out, sizeof(double) * 2)) return fail(SCUFFX_ERR_ALLOC, "allocation failed");
  return 0;
}

int scuffx_rhs_planewave(scuffx_geom *G, double wr, double wi, const double *E0, const double *nhat, void *rhs)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !E0 || !nhat || !rhs) return fail(SCUFFX_ERR_ARG, "null argument");
  Assembler *A = G->assembler();
  if (!A->error().empty()) return fail(SCUFFX_ERR_HIP, A->error());
  int rc = ensure_rhs_data(G);
  if (rc) return rc;
  cdouble omega(wr, wi), eps, mu;
  G->G.Regions[0].eps_mu(omega, &eps, &mu);
  cdouble K = std::sqrt(eps * mu) * omega, Z = 376.73031346177 * std::sqrt(mu / eps);
  cdouble e0[3] = {cdouble(E0[0], E0[1]), cdouble(E0[2], E0[3]), cdouble(E0[4], E0[5])};
  hipMemsetAsync(rhs, 0, sizeof(double2) * G->G.TotalBFs, g_stream);
  for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
    const Surface &S = G->G.Surfaces[s];
    double sgn;
    if (S.RegionIndices[0] == 0) sgn = -1.0;
    else if (S.RegionIndices[1] == 0) sgn = 1.0;
    else continue;
    const DevSurface &D = A->dev_surface((int)s);
    if (launch_rhs(g_stream, D.V, D.PVI, G->rhs[s].PA, G->rhs[s].EV, D.EI, D.EL, D.nE, S.IsPEC, sgn, e0, nhat, K, Z,
                   (double2 *)rhs + G->G.BFIndexOffset[s]) != hipSuccess)
      return fail(SCUFFX_ERR_HIP, "RHS kernel failed");
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "RHS kernel failed");
  return 0;
}

int scuffx_power(scuffx_geom *G, double wr, double wi, const void *KN, const void *rhs, double *PAbs, double *PExt)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  (void)wr; (void)wi;
  if (!G || !KN || !rhs) return fail(SCUFFX_ERR_ARG, "null argument");
  Assembler *A = G->assembler();
  int rc = ensure_rhs_data(G);
  if (rc) return rc;
  double tot[2] = {0, 0};
  for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
    const Surface &S = G->G.Surfaces[s];
    const DevSurface &D = A->dev_surface((int)s);
    int off = G->G.BFIndexOffset[s];
    if (launch_power(g_stream, D.V, D.PVI, G->rhs[s].PA, G->rhs[s].PZ, G->rhs[s].PEI, D.EI, D.EL, D.nE, S.IsPEC,
                     (const double2 *)KN + off, (const double2 *)rhs + off, G->rhs[s].work, G->d_out) != hipSuccess)
      return fail(SCUFFX_ERR_HIP, "power kernel failed");
    double h[2];
    hipMemcpyAsync(h, G->d_out, sizeof(h), hipMemcpyDeviceToHost, g_stream);
    hipStreamSynchronize(g_stream);
    tot[0] += h[0]; tot[1] += h[1];
  }
  if (PExt) *PExt = tot[0];
  if (PAbs) *PAbs = tot[1];
  return 0;
}

int scuffx_assemble_region(scuffx_geom *G, double wr, double wi, int region, void *M, long ldm)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !M || G->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  if (region < 0 || region >= (int)G->G.Regions.size()) return fail(SCUFFX_ERR_ARG, "region index out of range");
  Assembler *A = G->assembler();
  A->set_region_filter(region);
  int rc = A->assemble(cdouble(wr, wi), M, ldm, MAT_C128, g_stream);
  A->set_region_filter(-1);
  if (rc) return fail(rc, A->error());
  return 0;
}

int scuffx_power_emt(scuffx_geom *G, double wr, double wi, const void *KN, const void *rhs, const void *Mext,
                     long ldm, double *PAbs, double *PScat)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !KN || !rhs || !Mext) return fail(SCUFFX_ERR_ARG, "null argument");
  const int n = G->G.TotalBFs;
  if (ldm < n) return fail(SCUFFX_ERR_ARG, "leading dimension smaller than the matrix");
  for (const Surface &S : G->G.Surfaces)
    if (S.RegionIndices[0] != 0) return fail(SCUFFX_ERR_UNSUPPORTED, "EMT power needs every surface in the exterior region");
  double pext = 0.0, pabs_opft = 0.0;
  int rc = scuffx_power(G, wr, wi, KN, rhs, &pabs_opft, &pext);
  if (rc) return rc;
  if (!G->d_bft) {
    std::vector<signed char> t(n);
    for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
      const Surface &S = G->G.Surfaces[s];
      const int off = G->G.BFIndexOffset[s];
      for (int e = 0; e < S.NumEdges(); e++) {
        if (S.IsPEC) t[off + e] = 0;
        else { t[off + 2 * e] = 1; t[off + 2 * e + 1] = 2; }
      }
    }
    if (hipMalloc(&G->d_bft, n) != hipSuccess || hipMalloc(&G->d_emt, sizeof(double) * (n + 1)) != hipSuccess)
      return fail(SCUFFX_ERR_ALLOC, "allocation failed");
    hipMemcpy(G->d_bft, t.data(), n, hipMemcpyHostToDevice);
  }
  if (launch_emt(g_stream, (const double2 *)Mext, ldm, (const double2 *)KN, G->d_bft, n, G->d_emt, G->d_emt + n) !=
      hipSuccess)
    return fail(SCUFFX_ERR_HIP, "EMT power kernel failed");
  double ps = 0.0;
  hipMemcpyAsync(&ps, G->d_emt + n, sizeof(double), hipMemcpyDeviceToHost, g_stream);
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "EMT power kernel failed");
  if (PScat) *PScat = ps;
  if (PAbs) *PAbs = pext - ps;
  return 0;
}

int scuffx_vertex_slots(scuffx_geom *G)
{
  if (!G) return fail(SCUFFX_ERR_ARG, "null geometry");
  long n = 0;
  for (const Surface &S : G->G.Surfaces) n += 3L * S.NumVertices();
  return (int)n;
}

static int vg_code(int rc) { return rc == -2 ? SCUFFX_ERR_ALLOC : rc == -4 ? SCUFFX_ERR_UNSUPPORTED : SCUFFX_ERR_HIP; }

int scuffx_vertex_gradient(scuffx_geom *G, double wr, double wi, const void *M, long ldm, const void *u,
                           const void *v, const void *C, long ldc, void *grad)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !M || !grad || (!C && (!u || !v))) return fail(SCUFFX_ERR_ARG, "bad arguments");
  const long n = G->G.TotalBFs;
  if (ldm < n || (C && ldc < n)) return fail(SCUFFX_ERR_ARG, "leading dimension below the matrix order");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  Assembler *A = G->assembler();
  if (!A->error().empty()) return fail(SCUFFX_ERR_HIP, A->error());
  VGW w;
  w.u = (const double2 *)u; w.v = (const double2 *)v; w.C = (const double2 *)C; w.ldc = ldc;
  std::string msg;
  const int rc = A->vertex_gradient(cdouble(wr, wi), (const double2 *)M, ldm, w, (double2 *)grad, g_stream, &msg);
  return rc ? fail(vg_code(rc), msg) : 0;
}

int scuffx_rhs_vertex_gradient(scuffx_geom *G, double wr, double wi, const double *E0, const double *nhat,
                               const void *mu, void *grad)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !E0 || !nhat || !mu || !grad) return fail(SCUFFX_ERR_ARG, "null argument");
  Assembler *A = G->assembler();
  if (!A->error().empty()) return fail(SCUFFX_ERR_HIP, A->error());
  int rc = ensure_rhs_data(G);
  if (rc) return rc;
  if (A->vg_tables()) return fail(SCUFFX_ERR_ALLOC, "allocation failed");
  cdouble omega(wr, wi), eps, mu0;
  G->G.Regions[0].eps_mu(omega, &eps, &mu0);
  const cdouble K = std::sqrt(eps * mu0) * omega, Z = 376.73031346177 * std::sqrt(mu0 / eps);
  const cdouble e0[3] = {cdouble(E0[0], E0[1]), cdouble(E0[2], E0[3]), cdouble(E0[4], E0[5])};
  if (hipMemsetAsync(grad, 0, sizeof(double2) * A->vg_slots(), g_stream) != hipSuccess)
    return fail(SCUFFX_ERR_HIP, "HIP error");
  for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
    const Surface &S = G->G.Surfaces[s];
    double sgn;
    if (S.RegionIndices[0] == 0) sgn = -1.0;
    else if (S.RegionIndices[1] == 0) sgn = 1.0;
    else continue;
    if (launch_vg_rhs(g_stream, A->vg_surface((int)s), A->vg_edges((int)s), S.NumEdges(), sgn, e0, nhat, K, Z,
                      (const double2 *)mu, (double2 *)grad) != hipSuccess)
      return fail(SCUFFX_ERR_HIP, "RHS gradient kernel failed");
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "RHS gradient kernel failed");
  return 0;
}

int scuffx_pabs_vertex_gradient(scuffx_geom *G, const void *KN, double *grad, void *gamma)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !KN || !grad || !gamma) return fail(SCUFFX_ERR_ARG, "null argument");
  Assembler *A = G->assembler();
  if (!A->error().empty()) return fail(SCUFFX_ERR_HIP, A->error());
  int rc = ensure_rhs_data(G);
  if (rc) return rc;
  if (A->vg_tables()) return fail(SCUFFX_ERR_ALLOC, "allocation failed");
  if (hipMemsetAsync(grad, 0, sizeof(double) * A->vg_slots(), g_stream) != hipSuccess ||
      hipMemsetAsync(gamma, 0, sizeof(double2) * G->G.TotalBFs, g_stream) != hipSuccess)
    return fail(SCUFFX_ERR_HIP, "HIP error");
  for (size_t s = 0; s < G->G.Surfaces.size(); s++) {
    const Surface &S = G->G.Surfaces[s];
    if (S.IsPEC) continue;
    const DevSurface &D = A->dev_surface((int)s);
    if (launch_vg_pabs(g_stream, A->vg_surface((int)s), G->rhs[s].PZ, G->rhs[s].PEI, D.EI, D.nE,
                       (const double2 *)KN, grad, (double2 *)gamma) != hipSuccess)
      return fail(SCUFFX_ERR_HIP, "power gradient kernel failed");
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) return fail(SCUFFX_ERR_HIP, "power gradient kernel failed");
  return 0;
}

int scuffx_adjoint(scuffx_geom *G, int ns, double wr, double wi, const void *x, const void *y, int ntorque,
                   const double *gamma, double *out)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !x || !y || !out || ns < 0 || ns >= (int)G->G.Surfaces.size() || ntorque < 0 || ntorque > 3 ||
      (ntorque > 0 && !gamma))
    return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (wr == 0.0 && wi == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  const int NS = (int)G->G.Surfaces.size(), nq = 3 + ntorque, nparts = 64;
  const int na = G->G.Surfaces[ns].NumBFs, oa = G->G.BFIndexOffset[ns];
  int nbmax = 0;
  for (int b = 0; b < NS; b++)
    if (b != ns) nbmax = std::max(nbmax, G->G.Surfaces[b].NumBFs);
  for (int m = 0; m < 2 * nq; m++) out[m] = 0;
  if (nbmax == 0) return 0;
  Assembler *A = G->assembler();
  const size_t blk = (size_t)na * nbmax;
  double2 *B = nullptr, *GB = nullptr, *TB = nullptr, *part = nullptr;
  int rc = 0;
  if (hipMalloc(&B, sizeof(double2) * blk) || hipMalloc(&GB, sizeof(double2) * 3 * blk) ||
      (ntorque && hipMalloc(&TB, sizeof(double2) * ntorque * blk)) ||
      hipMalloc(&part, sizeof(double2) * nparts * nq * NS))
    rc = fail(SCUFFX_ERR_ALLOC, "allocation failed");
  const double2 *X = (const double2 *)x, *Y = (const double2 *)y;
  for (int b = 0; b < NS && rc == 0; b++) {
    if (b == ns) continue;
    const int nb = G->G.Surfaces[b].NumBFs, ob = G->G.BFIndexOffset[b];
    const long st = (long)na * nb;
    int r = A->assemble_block(ns, b, cdouble(wr, wi), B, na, GB, na, st, ntorque, gamma, TB, g_stream);
    if (r) { rc = fail(r, A->error()); break; }
    for (int m = 0; m < nq; m++) {
      const double2 *Gm = m < 3 ? GB + m * st : TB + (m - 3) * st;
      if (launch_adj_contract(g_stream, Gm, na, na, nb, X + oa, Y + oa, X + ob, Y + ob,
                              part + ((long)b * nq + m) * nparts, nparts) != hipSuccess)
        rc = fail(SCUFFX_ERR_HIP, "contraction kernel failed");
    }
  }
  if (rc == 0) {
    std::vector<double2> h((size_t)nparts * nq * NS);
    hipMemcpyAsync(h.data(), part, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, g_stream);
    if (hipStreamSynchronize(g_stream) != hipSuccess) rc = fail(SCUFFX_ERR_HIP, "contraction failed");
    for (int b = 0; b < NS && rc == 0; b++) {
      if (b == ns) continue;
      for (int m = 0; m < nq; m++)
        for (int k = 0; k < nparts; k++) {
          out[2 * m] += h[((size_t)b * nq + m) * nparts + k].x;
          out[2 * m + 1] += h[((size_t)b * nq + m) * nparts + k].y;
        }
    }
  }
  hipFree(B); hipFree(GB); hipFree(TB); hipFree(part);
  return rc;
}

int scuffx_batch_sweep(scuffx_geom *G, const double *omegas, int count, const double *E0, const double *nhat,
                       int prec, double *out)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!G || !omegas || count < 0 || !E0 || !nhat || !out || G->G.Surfaces.empty())
    return fail(SCUFFX_ERR_ARG, "bad arguments");
  if (prec != 64 && prec != 32) return fail(SCUFFX_ERR_ARG, "prec must be 64 or 32");
  const long n = G->G.TotalBFs;
  double2 *M = nullptr, *rhs = nullptr, *kn = nullptr;
  float2 *lu32 = nullptr;
  int *ipiv = nullptr;
  int rc = 0;
  if (hipMalloc(&M, sizeof(double2) * n * n) || hipMalloc(&rhs, sizeof(double2) * n) ||
      hipMalloc(&kn, sizeof(double2) * n) || hipMalloc(&ipiv, sizeof(int) * n) ||
      (prec == 32 && hipMalloc(&lu32, sizeof(float2) * n * n)))
    rc = fail(SCUFFX_ERR_ALLOC, "allocation failed");
  for (int k = 0; k < count && rc == 0; k++) {
    const double wr = omegas[2 * k], wi = omegas[2 * k + 1];
    if ((rc = scuffx_assemble(G, wr, wi, M, n)) != 0) break;
    if ((rc = scuffx_rhs_planewave(G, wr, wi, E0, nhat, rhs)) != 0) break;
    if (prec == 64) {
      hipMemcpyAsync(kn, rhs, sizeof(double2) * n, hipMemcpyDeviceToDevice, g_stream);
      if ((rc = scuffx_lu_factor(M, n, n, ipiv, 64)) != 0) break;
      if ((rc = scuffx_lu_solve(M, n, n, ipiv, 'N', 1, kn, n, 64)) != 0) break;
    } else {
      if ((rc = scuffx_lu32_from_c128(M, n, n, lu32, n, ipiv)) != 0) break;
      int it = 0;
      double rel = 0;
      if ((rc = scuffx_refine(M, n, n, lu32, n, ipiv, rhs, kn, 1e-13, 30, &it, &rel)) != 0) {
        if (rc == SCUFFX_NOT_CONVERGED) rc = fail(SCUFFX_ERR_NUMERIC, g_err);
        break;
      }
    }
    double pa = 0, pe = 0;
    if ((rc = scuffx_power(G, wr, wi, kn, rhs, &pa, &pe)) != 0) break;
    out[3 * k] = pa;
    out[3 * k + 1] = pe;
    out[3 * k + 2] = pe - pa;
  }
  hipStreamSynchronize(g_stream);
  hipFree(M); hipFree(rhs); hipFree(kn); hipFree(ipiv); hipFree(lu32);
  return rc;
}

// ---------------------------------------------------------------------
// Batched frequency sweep over several geometries (uneven N): for each
// geometry every frequency's matrix is assembled, the whole batch is
// converted to complex64 and factored by the batched tournament LU, the
// plane-wave RHS is solved and refined in complex128 with a fixed number
// of steps, and the power is computed -- all recorded once into one
// hipGraph by stream capture and launched once.  The frequency-independent
// FIPPI table and geometry are built (or taken from the cache) first.
// Instances whose factorization reports a problem or whose refined
// residual stays above 1e-11 are solved again by the step-by-step mixed
// path (classical refinement continued by GMRES-IR).
// ---------------------------------------------------------------------
}  // extern "C"

namespace {
struct SweepBuf {
  long n = 0, NS = 0;
  int nfreq = 0;
  double2 *M = nullptr, *rhs = nullptr, *x = nullptr, *r = nullptr, *work = nullptr;
  float2 *L = nullptr, *x32 = nullptr;
  int *ipiv = nullptr, *status = nullptr;
  void *scratch = nullptr, *sws = nullptr;
  double *outs = nullptr, *relres = nullptr, *kz = nullptr;
  int nbmax = 0;
  void release()
  {
    if (hip_down()) return;
    hipFree(M); hipFree(rhs); hipFree(x); hipFree(r); hipFree(work); hipFree(L); hipFree(x32); hipFree(ipiv);
    hipFree(status); hipFree(scratch); hipFree(sws); hipFree(outs); hipFree(relres); hipFree(kz);
    *this = SweepBuf();
  }
  int alloc(long n_, long NS_, int nf)
  {
    if (M && n == n_ && NS == NS_ && nfreq >= nf) return 0;
    release();
    n = n_; NS = NS_; nfreq = nf;
    nbmax = std::min(nf, lu_batch_max(n));
    const size_t nn = (size_t)n * n, nch = (size_t)residual_chunks(n);
    if (nbmax <= 0 || hipMalloc(&M, sizeof(double2) * nn * nf) || hipMalloc(&L, sizeof(float2) * nn * nf) ||
        hipMalloc(&rhs, sizeof(double2) * n * nf) || hipMalloc(&x, sizeof(double2) * n * nf) ||
        hipMalloc(&r, sizeof(double2) * n * nf) || hipMalloc(&x32, sizeof(float2) * n * nf) ||
        hipMalloc(&work, sizeof(double2) * n * nch * nf) || hipMalloc(&ipiv, sizeof(int) * n * nf) ||
        hipMalloc(&status, sizeof(int) * 2 * nf) || hipMalloc(&scratch, lu_batch_scratch_bytes(n) * nbmax) ||
        hipMalloc(&sws, lu_solve_batch_bytes(n) * nf) || hipMalloc(&outs, sizeof(double) * 2 * NS * nf) ||
        hipMalloc(&relres, sizeof(double) * nf) || hipMalloc(&kz, sizeof(double) * 4 * nf)) {
      release();
      return -2;
    }
    return 0;
  }
};

// The recorded sweep of one host thread, launched again by the next sweep
// of the same shape: same geometries (assembler objects and their buffer
// generations), orders, frequency count, plane wave, refinement steps and
// region counts.  Frequency-dependent values reach it through device
// tables (region constants, K and Z of the exterior medium).
struct SweepGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned long long> key;
  std::vector<double> fkey;
  void release()
  {
    if (exec && !hip_down()) hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
    fkey.clear();
  }
};
struct SweepState {
  std::vector<SweepBuf> bufs;
  SweepGraph graph;
};

// the sweep buffers and graph of each host thread, kept for the next sweep
// (released by scuffx_shutdown)
std::mutex g_sweep_mu;
std::vector<SweepState *> g_sweep_all;
SweepState &sweep_state()
{
  static thread_local SweepState *v = nullptr;
  if (!v) {
    v = new SweepState();
    std::lock_guard<std::mutex> lk(g_sweep_mu);
    g_sweep_all.push_back(v);
  }
  return *v;
}
void sweep_release_all()
{
  std::lock_guard<std::mutex> lk(g_sweep_mu);
  for (auto *v : g_sweep_all) {
    v->graph.release();
    for (auto &b : v->bufs) b.release();
  }
}

// autotuned block width of the batched LU per (order, batch) class: the
// faster of 128 and 256 columns, timed once per process on the first
// batch of that class (the trial factors are overwritten afterwards)
std::mutex g_tune_mu;
std::vector<std::pair<std::pair<long, int>, int>> g_tuned;

bool tuned_lookup(long n, int nbat, int *nb)
{
  const std::pair<long, int> key{(n + 255) / 256, nbat};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto &e : g_tuned)
    if (e.first == key) { *nb = e.second; return true; }
  return false;
}

int tune_nb(long n, int nbat, SweepBuf &B, hipStream_t s)
{
  if (const char *e = getenv("SCUFFX_SWEEP_NB")) return std::max(32, std::min(256, atoi(e) / 32 * 32));
  int best = 256;
  float best_ms = 1e30f;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  for (int nb : {128, 256}) {
    convert_c128_to_c64(B.M, B.L, n, (long)nbat * n, n, n, s);
    hipEventRecord(e0, s);
    int rc = lu_factor_batch(B.L, n, n, nbat, n * n, B.ipiv, n, B.scratch, B.status, s, nb);
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    if (rc == 0 && ms < best_ms) { best_ms = ms; best = nb; }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.push_back({{(n + 255) / 256, nbat}, best});
  return best;
}

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

extern "C" {

int scuffx_sweep_batched(scuffx_geom **Gs, int ngeom, const double *omegas, int nfreq, const double *E0,
                         const double *nhat, int refine_steps, double *out, double *relres, int *info, double *ms)
{
  if (hip_down()) return fail(SCUFFX_ERR_ARG, "library shut down");
  if (!Gs || ngeom <= 0 || !omegas || nfreq <= 0 || !E0 || !nhat || !out || refine_steps < 0 || refine_steps > 20)
    return fail(SCUFFX_ERR_ARG, "bad arguments");
  for (int g = 0; g < ngeom; g++)
    if (!Gs[g] || Gs[g]->G.Surfaces.empty()) return fail(SCUFFX_ERR_ARG, "empty geometry");
  std::vector<cdouble> om(nfreq);
  for (int f = 0; f < nfreq; f++) {
    om[f] = cdouble(omegas[2 * f], omegas[2 * f + 1]);
    if (om[f] == 0.0) return fail(SCUFFX_ERR_ARG, "zero frequency");
  }
  static thread_local hipStream_t cs = nullptr;   // capture stream (default priority)
  if (!cs && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return fail(SCUFFX_ERR_HIP, "stream creation failed");
  static bool warmed[64] = {false};
  int dev = 0;
  hipGetDevice(&dev);
  const cdouble e0[3] = {cdouble(E0[0], E0[1]), cdouble(E0[2], E0[3]), cdouble(E0[4], E0[5])};
  double t[6] = {now_ms(), 0, 0, 0, 0, 0};
  SweepState &state = sweep_state();
  std::vector<SweepBuf> &bufs = state.bufs;
  if ((int)bufs.size() < ngeom) bufs.resize(ngeom);
  int rc = 0;
  std::vector<int> nbsel(ngeom, 256);
  // ---- set-up outside the capture: tables, buffers, first-use state ----
  for (int g = 0; g < ngeom && rc == 0; g++) {
    scuffx_geom *G = Gs[g];
    SweepBuf &B = bufs[g];
    const long n = G->G.TotalBFs, NS = (long)G->G.Surfaces.size();
    Assembler *A = G->assembler();
    if ((rc = A->prepare_graph(om, cs)) != 0) { rc = fail(rc, A->error()); break; }
    if ((rc = ensure_rhs_data(G)) != 0) break;
    if (n > 20480 || B.alloc(n, NS, nfreq) != 0) {
      rc = n > 20480 ? fail(SCUFFX_ERR_UNSUPPORTED, "order too large for the batched LU")
                     : fail(SCUFFX_ERR_ALLOC, "allocation failed");
      break;
    }
    const bool tuned = tuned_lookup(n, B.nbmax, &nbsel[g]);
    if (!warmed[dev & 63] || !tuned) {
      // first use: the kernels' constant tables (an ordinary assembly and
      // RHS of the first frequency) and the LU block width
      if ((rc = A->assemble(om[0], B.M, n, MAT_C128, cs)) != 0) { rc = fail(rc, A->error()); break; }
      cdouble eps, mu;
      G->G.Regions[0].eps_mu(om[0], &eps, &mu);
      cdouble K = std::sqrt(eps * mu) * om[0], Z = 376.73031346177 * std::sqrt(mu / eps);
      const DevSurface &D = A->dev_surface(0);
      if (launch_rhs(cs, D.V, D.PVI, G->rhs[0].PA, G->rhs[0].EV, D.EI, D.EL, D.nE, G->G.Surfaces[0].IsPEC, -1.0, e0,
                     nhat, K, Z, B.rhs) != hipSuccess) { rc = fail(SCUFFX_ERR_HIP, "RHS kernel failed"); break; }
      warmed[dev & 63] = true;
      if (!tuned) {
        const size_t nn = (size_t)n * n;
        for (int f = 1; f < B.nbmax; f++)
          hipMemcpyAsync(B.M + nn * f, B.M, sizeof(double2) * nn, hipMemcpyDeviceToDevice, cs);
        nbsel[g] = tune_nb(n, B.nbmax, B, cs);
      }
    }
    if (hipStreamSynchronize(cs) != hipSuccess) { rc = fail(SCUFFX_ERR_HIP, "sweep set-up failed"); break; }
  }
  if (rc) { hipStreamSynchronize(cs); return rc; }
  // K and Z of the exterior medium per frequency, read by the recorded RHS
  std::vector<unsigned long long> key = {(unsigned long long)ngeom, (unsigned long long)nfreq,
                                         (unsigned long long)refine_steps, (unsigned long long)dev};
  for (int g = 0; g < ngeom; g++) {
    scuffx_geom *G = Gs[g];
    Assembler *A = G->assembler();
    std::vector<double> kz(4 * nfreq);
    for (int f = 0; f < nfreq; f++) {
      cdouble eps, mu;
      G->G.Regions[0].eps_mu(om[f], &eps, &mu);
      cdouble K = std::sqrt(eps * mu) * om[f], Z = 376.73031346177 * std::sqrt(mu / eps);
      kz[4 * f] = K.real(); kz[4 * f + 1] = K.imag(); kz[4 * f + 2] = Z.real(); kz[4 * f + 3] = Z.imag();
    }
    if (hipMemcpy(bufs[g].kz, kz.data(), sizeof(double) * 4 * nfreq, hipMemcpyHostToDevice) != hipSuccess)
      return fail(SCUFFX_ERR_HIP, "sweep set-up failed");
    key.push_back(A->serial());
    key.push_back(A->alloc_generation());
    key.push_back((unsigned long long)(uintptr_t)bufs[g].M);
    key.push_back((unsigned long long)nbsel[g]);
    key.push_back((unsigned long long)A->far_precision());
    for (int v : A->nreg_table()) key.push_back((unsigned long long)v);
  }
  std::vector<double> fkey(E0, E0 + 6);
  fkey.insert(fkey.end(), nhat, nhat + 3);
  if (hipStreamSynchronize(cs) != hipSuccess) return fail(SCUFFX_ERR_HIP, "sweep set-up failed");
  t[1] = now_ms();
  const bool reuse = state.graph.exec && state.graph.key == key && state.graph.fkey == fkey;
  int crc = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = reuse ? state.graph.exec : nullptr;
  if (!reuse) {
  state.graph.release();
  // ---- the sweep, recorded into one graph ----
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(SCUFFX_ERR_HIP, "stream capture failed to start");
  for (int g = 0; g < ngeom && crc == 0; g++) {
    scuffx_geom *G = Gs[g];
    SweepBuf &B = bufs[g];
    Assembler *A = G->assembler();
    const long n = B.n, nn = n * n, NS = B.NS;
    for (int f = 0; f < nfreq && crc == 0; f++) crc = A->assemble_graph(f, B.M + nn * f, n, cs);
    if (crc) break;
    convert_c128_to_c64(B.M, B.L, n, (long)nfreq * n, n, n, cs);
    for (int f0 = 0; f0 < nfreq && crc == 0; f0 += B.nbmax) {
      const int nb = std::min(B.nbmax, nfreq - f0);
      crc = lu_factor_batch(B.L + nn * f0, n, n, nb, nn, B.ipiv + n * f0, n, B.scratch, B.status + 2 * f0, cs, nbsel[g]);
      if (!crc) crc = lu_solve_batch_prepare(B.L + nn * f0, n, n, nb, nn, B.ipiv + n * f0, n,
                                             (char *)B.sws + lu_solve_batch_bytes(n) * f0, cs);
    }
    if (crc) break;
    for (int f = 0; f < nfreq; f++) {
      zero_c128(B.rhs + n * f, n, cs);
      for (long s = 0; s < NS; s++) {
        const Surface &S = G->G.Surfaces[s];
        double sgn;
        if (S.RegionIndices[0] == 0) sgn = -1.0;
        else if (S.RegionIndices[1] == 0) sgn = 1.0;
        else continue;
        const DevSurface &D = A->dev_surface((int)s);
        launch_rhs(cs, D.V, D.PVI, G->rhs[s].PA, G->rhs[s].EV, D.EI, D.EL, D.nE, S.IsPEC, sgn, e0, nhat, 0.0, 1.0,
                   B.rhs + n * f + G->G.BFIndexOffset[s], B.kz + 4 * f);
      }
    }
    auto solve32 = [&](float2 *v) {
      for (int f0 = 0; f0 < nfreq; f0 += B.nbmax) {
        const int nb = std::min(B.nbmax, nfreq - f0);
        lu_solve_batch(B.L + nn * f0, n, n, nb, nn, v + n * f0, n, (char *)B.sws + lu_solve_batch_bytes(n) * f0, cs);
      }
    };
    // x0 = LU32^{-1} b, then x += LU32^{-1} (b - M x) in complex128
    convert_c128_to_c64(B.rhs, B.x32, n, nfreq, n, n, cs);
    solve32(B.x32);
    convert_c64_to_c128(B.x32, B.x, n * nfreq, cs);
    for (int it = 0; it < refine_steps; it++) {
      residual_c128_batch(B.M, n, n, nn, nfreq, B.x, B.rhs, B.r, B.work, cs);
      convert_c128_to_c64(B.r, B.x32, n, nfreq, n, n, cs);
      solve32(B.x32);
      axpy_c64_into_c128(B.x32, B.x, n * nfreq, cs);
    }
    residual_c128_batch(B.M, n, n, nn, nfreq, B.x, B.rhs, B.r, B.work, cs);
    batch_relres(B.r, B.rhs, n, nfreq, B.relres, cs);
    for (int f = 0; f < nfreq; f++)
      for (long s = 0; s < NS; s++) {
        const Surface &S = G->G.Surfaces[s];
        const DevSurface &D = A->dev_surface((int)s);
        const long off = G->G.BFIndexOffset[s];
        launch_power(cs, D.V, D.PVI, G->rhs[s].PA, G->rhs[s].PZ, G->rhs[s].PEI, D.EI, D.EL, D.nE, S.IsPEC,
                     B.x + n * f + off, B.rhs + n * f + off, G->rhs[s].work, B.outs + 2 * (f * NS + s));
      }
  }
  hipError_t ce = hipStreamEndCapture(cs, &graph);
  if (crc || ce != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    hipStreamSynchronize(cs);
    return crc ? fail(crc, "sweep capture failed") : fail(SCUFFX_ERR_HIP, "stream capture failed");
  }
  t[2] = now_ms();
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(SCUFFX_ERR_HIP, "graph instantiation failed");
  state.graph.exec = exec;
  state.graph.key = key;
  state.graph.fkey = fkey;
  } else {
    t[2] = now_ms();
  }
  t[3] = now_ms();
  hipError_t le = hipGraphLaunch(exec, cs);
  hipError_t se = hipStreamSynchronize(cs);
  t[4] = now_ms();
  if (le != hipSuccess || se != hipSuccess) {
    state.graph.release();
    return fail(SCUFFX_ERR_HIP, "sweep graph failed on the device");
  }
  // ---- results; instances that need it are solved again step by step ----
  int redone = 0;
  for (int g = 0; g < ngeom && rc == 0; g++) {
    scuffx_geom *G = Gs[g];
    SweepBuf &B = bufs[g];
    const long n = B.n, nn = n * n, NS = B.NS;
    std::vector<int> st(2 * nfreq);
    std::vector<double> rel(nfreq), po(2 * NS * nfreq);
    hipMemcpy(st.data(), B.status, sizeof(int) * 2 * nfreq, hipMemcpyDeviceToHost);
    hipMemcpy(rel.data(), B.relres, sizeof(double) * nfreq, hipMemcpyDeviceToHost);
    hipMemcpy(po.data(), B.outs, sizeof(double) * 2 * NS * nfreq, hipMemcpyDeviceToHost);
    for (int f = 0; f < nfreq && rc == 0; f++) {
      double pe = 0, pa = 0;
      for (long s = 0; s < NS; s++) { pe += po[2 * (f * NS + s)]; pa += po[2 * (f * NS + s) + 1]; }
      double rr = rel[f];
      if (st[2 * f] != 0 || st[2 * f + 1] != 0 || !(rel[f] <= 1e-11)) {
        // the step-by-step mixed path on this instance's complex128 matrix
        std::swap(g_stream, cs);
        int it = 0;
        double r2 = 0;
        int rc2 = scuffx_lu32_from_c128(B.M + nn * f, n, n, B.L + nn * f, n, B.ipiv + n * f);
        if (rc2 == 0)
          rc2 = scuffx_refine(B.M + nn * f, n, n, B.L + nn * f, n, B.ipiv + n * f, B.rhs + n * f, B.x + n * f, 1e-13,
                              30, &it, &r2);
        if (rc2 == 0 || rc2 == SCUFFX_NOT_CONVERGED) rc2 = scuffx_power(G, om[f].real(), om[f].imag(), B.x + n * f,
                                                                       B.rhs + n * f, &pa, &pe);
        std::swap(g_stream, cs);
        if (rc2 < 0) { rc = rc2; break; }
        rr = r2;
        redone++;
      }
      double *o = out + 3 * ((long)g * nfreq + f);
      o[0] = pa; o[1] = pe; o[2] = pe - pa;
      if (relres) relres[(long)g * nfreq + f] = rr;
    }
  }
  t[5] = now_ms();
  if (info) { info[0] = redone; info[1] = nbsel[0]; }
  if (ms) for (int k = 0; k < 5; k++) ms[k] = t[k + 1] - t[k];
  return rc;
}

int scuffx_profile(int on)
{
  prof_collect();
  prof_reset();
  prof_enable(on != 0);
  return 0;
}

int scuffx_profile_read(const char *kernel, double *ms, double *work, long *launches)
{
  if (!kernel) return fail(SCUFFX_ERR_ARG, "null kernel name");
  prof_collect();
  for (int k = 0; k < PK_COUNT; k++)
    if (!std::strcmp(kernel, prof_name(k))) { prof_read(k, ms, work, launches); return 0; }
  return fail(SCUFFX_ERR_ARG, std::string("unknown kernel name ") + kernel);
}

}  // extern "C"
