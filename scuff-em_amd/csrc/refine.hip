// refine.hip -- complex128 solutions from complex64 LU factors.
//
// Two refinements of x ~ M^{-1} b, both with the residual b - M x in
// complex128 (one pass over M, 16 n^2 bytes):
//  * classical iterative refinement: x += LU32^{-1} (b - M x) while the
//    residual keeps falling; the best iterate is kept;
//  * GMRES-IR: each correction solves M d = r with restarted flexible
//    GMRES in complex128, right-preconditioned by the complex64 factors
//    (Carson and Higham, SIAM J. Sci. Comput. 40 (2018) A817).  Arnoldi
//    uses classical Gram-Schmidt applied twice, the small Hessenberg
//    least-squares problem is solved on the host with Givens rotations.
// The reference has no mixed-precision path (its drop-in is the plain
// complex128 LU, LBWrappers.cc:151-245); these only change how the
// complex128 answer is reached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>

#include "lu.h"
#include "prof.h"
#include "refine.h"

namespace scuffx {

static size_t r256(size_t b) { return 256 * ((b + 255) / 256); }

size_t refine_ws_bytes(long n, int m)
{
  return 3 * r256(sizeof(double2) * n) + r256(sizeof(double2) * n * residual_chunks(n)) +
         r256(sizeof(float2) * n) + 256 + r256(sizeof(double2) * REFINE_PARTS * (REFINE_GMRES_MAX + 1)) +
         r256(sizeof(double2) * n * (2 * (long)m + 1)) + solve_ws_bytes<float2>(n);
}

RefineWs refine_ws_carve(char *base, long n, int m)
{
  RefineWs w;
  w.m = m;
  w.r = (double2 *)base; base += r256(sizeof(double2) * n);
  w.xbest = (double2 *)base; base += r256(sizeof(double2) * n);
  w.d = (double2 *)base; base += r256(sizeof(double2) * n);
  w.work = (double2 *)base; base += r256(sizeof(double2) * n * residual_chunks(n));
  w.r32 = (float2 *)base; base += r256(sizeof(float2) * n);
  w.scal = (double *)base; base += 256;
  w.part = (double2 *)base; base += r256(sizeof(double2) * REFINE_PARTS * (REFINE_GMRES_MAX + 1));
  w.V = (double2 *)base; base += r256(sizeof(double2) * n * ((long)m + 1));
  w.Z = (double2 *)base; base += r256(sizeof(double2) * n * (long)m);
  w.sws = solve_ws_carve<float2>(base, n);
  return w;
}

// h[j] partial sums: sum_i conj(V[i + j ldv]) w[i] over a strided share
// of the rows, REFINE_PARTS shares per column j = blockIdx.y
__global__ void __launch_bounds__(256)
k_mdot(const double2 *V, long ldv, const double2 *w, long n, double2 *part)
{
  __shared__ double sr[256], si[256];
  const double2 *v = V + (long)blockIdx.y * ldv;
  double ar = 0, ai = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double2 a = v[i], b = w[i];
    ar += a.x * b.x + a.y * b.y;
    ai += a.x * b.y - a.y * b.x;
  }
  sr[threadIdx.x] = ar; si[threadIdx.x] = ai;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if (threadIdx.x < q) { sr[threadIdx.x] += sr[threadIdx.x + q]; si[threadIdx.x] += si[threadIdx.x + q]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = make_double2(sr[0], si[0]);
}

struct Coef { double2 c[REFINE_GMRES_MAX + 1]; };

// y[i] += sum_j c[j] V[i + j ldv]  (j < k)
__global__ void __launch_bounds__(256)
k_mupdate(const double2 *V, long ldv, int k, Coef c, double2 *y, long n)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double2 acc = y[i];
  for (int j = 0; j < k; j++) {
    const double2 v = V[i + (long)j * ldv], a = c.c[j];
    acc.x += a.x * v.x - a.y * v.y;
    acc.y += a.x * v.y + a.y * v.x;
  }
  y[i] = acc;
}

__global__ void k_scale(double2 *x, long n, double s)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) { x[i].x *= s; x[i].y *= s; }
}

typedef std::complex<double> zc;

// h[j] = V_j^H w for j < k (host result)
static int mdot(const double2 *V, long ldv, int k, const double2 *w, long n, double2 *part, zc *h, hipStream_t s)
{
  if (k > REFINE_GMRES_MAX + 1) return -1;
  hipLaunchKernelGGL(k_mdot, dim3(REFINE_PARTS, (unsigned)k), dim3(256), 0, s, V, ldv, w, n, part);
  double2 hp[REFINE_PARTS * (REFINE_GMRES_MAX + 1)];
  if (hipMemcpyAsync(hp, part, sizeof(double2) * REFINE_PARTS * k, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -3;
  for (int j = 0; j < k; j++) {
    double a = 0, b = 0;
    for (int p = 0; p < REFINE_PARTS; p++) { a += hp[(size_t)j * REFINE_PARTS + p].x; b += hp[(size_t)j * REFINE_PARTS + p].y; }
    h[j] = zc(a, b);
  }
  return 0;
}

static void mupdate(const double2 *V, long ldv, int k, const zc *c, double sign, double2 *y, long n, hipStream_t s)
{
  Coef cf;
  for (int j = 0; j < REFINE_GMRES_MAX + 1; j++) cf.c[j] = j < k ? make_double2(sign * c[j].real(), sign * c[j].imag()) : make_double2(0, 0);
  hipLaunchKernelGGL(k_mupdate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, V, ldv, k, cf, y, n);
}

// x = LU32^{-1} v in complex64, promoted to complex128
static int precond(const float2 *LU32, long n, long ld32, const int *ipiv, const double2 *v, double2 *out,
                   RefineWs &w, hipStream_t s)
{
  convert_c128_to_c64(v, w.r32, n, 1, n, n, s);
  int rc = lu_solve<float2>(LU32, n, ld32, ipiv, 'N', 1, w.r32, n, s, &w.sws);
  if (rc) return rc;
  convert_c64_to_c128(w.r32, out, n, s);
  return 0;
}

static double resid_rel(const double2 *M, long n, long ldm, const double2 *x, const double2 *b, double bn,
                        RefineWs &w, hipStream_t s)
{
  int tok = prof_begin(s, PK_RESID);
  residual_c128(M, n, ldm, x, b, w.r, w.work, s);
  prof_end(s, tok, 16.0 * (double)n * (double)n);
  return absmax_c128(w.r, n, w.scal, s) / bn;
}

// one restarted FGMRES(m) cycle for M d = r (r in w.r, rnorm2 = |r|_2):
// d in w.d; returns the number of inner steps or a negative error
static int fgmres_cycle(const double2 *M, long n, long ldm, const float2 *LU32, long ld32, const int *ipiv,
                        double rnorm2, RefineWs &w, double inner_tol, hipStream_t s)
{
  const int m = w.m;
  if (m < 1 || m > REFINE_GMRES_MAX) return -1;
  double2 *V = w.V, *Z = w.Z;
  hipMemcpyAsync(V, w.r, sizeof(double2) * n, hipMemcpyDeviceToDevice, s);
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, V, n, 1.0 / rnorm2);
  // host arrays of fixed size: nothing here allocates
  zc H[(REFINE_GMRES_MAX + 1) * REFINE_GMRES_MAX], g[REFINE_GMRES_MAX + 1], cs[REFINE_GMRES_MAX],
      sn[REFINE_GMRES_MAX];
  for (auto &v : H) v = 0;
  for (auto &v : g) v = 0;
  g[0] = rnorm2;
  int j = 0, rc;
  zc h[REFINE_GMRES_MAX + 1], h2[REFINE_GMRES_MAX + 1];
  for (; j < m;) {
    double2 *vj = V + (long)j * n, *zj = Z + (long)j * n, *vn = V + (long)(j + 1) * n;
    if ((rc = precond(LU32, n, ld32, ipiv, vj, zj, w, s)) != 0) return rc;
    residual_c128(M, n, ldm, zj, nullptr, vn, w.work, s);   // vn = M z_j
    // classical Gram-Schmidt, twice
    if ((rc = mdot(V, n, j + 1, vn, n, w.part, h, s)) != 0) return rc;
    mupdate(V, n, j + 1, h, -1.0, vn, n, s);
    if ((rc = mdot(V, n, j + 1, vn, n, w.part, h2, s)) != 0) return rc;
    mupdate(V, n, j + 1, h2, -1.0, vn, n, s);
    zc nn;
    if ((rc = mdot(vn, n, 1, vn, n, w.part, &nn, s)) != 0) return rc;
    const double hn = std::sqrt(std::max(0.0, nn.real()));
    for (int i = 0; i <= j; i++) H[(size_t)j * (m + 1) + i] = h[i] + h2[i];
    H[(size_t)j * (m + 1) + j + 1] = hn;
    if (hn > 0) hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vn, n, 1.0 / hn);
    // previous rotations on column j, then a new one zeroing H(j+1, j)
    zc *Hc = H + (size_t)j * (m + 1);
    for (int i = 0; i < j; i++) {
      const zc a = Hc[i], b = Hc[i + 1];
      Hc[i] = std::conj(cs[i]) * a + std::conj(sn[i]) * b;
      Hc[i + 1] = -sn[i] * a + cs[i] * b;
    }
    const double ra = std::abs(Hc[j]), rb = std::abs(Hc[j + 1]), den = std::hypot(ra, rb);
    if (den == 0) { cs[j] = 1; sn[j] = 0; }
    else { cs[j] = Hc[j] / den; sn[j] = Hc[j + 1] / den; }
    Hc[j] = std::conj(cs[j]) * Hc[j] + std::conj(sn[j]) * Hc[j + 1];
    Hc[j + 1] = 0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = std::conj(cs[j]) * g[j];
    j++;
    if (std::abs(g[j]) <= inner_tol * rnorm2 || hn == 0) break;
  }
  // y = H(0:j, 0:j)^{-1} g, d = Z y
  zc y[REFINE_GMRES_MAX];
  for (int i = j - 1; i >= 0; i--) {
    zc t = g[i];
    for (int q = i + 1; q < j; q++) t -= H[(size_t)q * (m + 1) + i] * y[q];
    y[i] = t / H[(size_t)i * (m + 1) + i];
  }
  hipMemsetAsync(w.d, 0, sizeof(double2) * n, s);
  mupdate(Z, n, j, y, 1.0, w.d, n, s);
  return j;
}

int gmres_ir(const double2 *M, long n, long ldm, const float2 *LU32, long ld32, const int *ipiv,
             const double2 *b, double2 *x, double tol, int maxiter, bool x_given, int *iters, double *relres,
             RefineWs &w, hipStream_t s)
{
  *iters = 0;
  const double bn = absmax_c128(b, n, w.scal, s);
  if (!(bn > 0)) {
    hipMemsetAsync(x, 0, sizeof(double2) * n, s);
    *relres = 0;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -3;
  }
  int rc;
  if (!x_given) {
    w.sws.inv_ready = false;
    if ((rc = precond(LU32, n, ld32, ipiv, b, x, w, s)) != 0) return rc;
  }
  double best = 1e300, rel = 0;
  int total = 0;
  for (int outer = 0;; outer++) {
    rel = resid_rel(M, n, ldm, x, b, bn, w, s);
    if (!std::isfinite(rel)) break;
    if (rel < best) { best = rel; hipMemcpyAsync(w.xbest, x, sizeof(double2) * n, hipMemcpyDeviceToDevice, s); }
    if (rel <= tol || total >= maxiter || outer > 2 * maxiter) break;
    zc nn;
    if ((rc = mdot(w.r, n, 1, w.r, n, w.part, &nn, s)) != 0) return rc;
    const double rn2 = std::sqrt(std::max(0.0, nn.real()));
    if (!(rn2 > 0)) break;
    // each cycle asks for ten times less than the remaining gap to tol
    const double inner = std::min(1e-2, std::max(1e-14, 0.1 * tol / rel));
    int steps = fgmres_cycle(M, n, ldm, LU32, ld32, ipiv, rn2, w, inner, s);
    if (steps < 0) return steps;
    total += steps;
    hipLaunchKernelGGL(k_mupdate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w.d, n, 1,
                       Coef{{make_double2(1, 0)}}, x, n);
  }
  if (!(best < 1e300)) return -4;   // no finite residual at any iterate
  if (!(rel <= best)) {
    hipMemcpyAsync(x, w.xbest, sizeof(double2) * n, hipMemcpyDeviceToDevice, s);
    rel = best;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return -3;
  *iters = total;
  *relres = rel;
  return rel <= tol ? 0 : 1;
}

int refine_ir(const double2 *M, long n, long ldm, const float2 *LU32, long ld32, const int *ipiv,
              const double2 *b, double2 *x, double tol, int maxiter, int gmres_fallback, int *iters,
              double *relres, RefineWs &w, hipStream_t s)
{
  *iters = 0;
  const double bn = absmax_c128(b, n, w.scal, s);
  if (!(bn > 0)) {
    hipMemsetAsync(x, 0, sizeof(double2) * n, s);
    *relres = 0;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -3;
  }
  w.sws.inv_ready = false;
  int rc = precond(LU32, n, ld32, ipiv, b, x, w, s);
  if (rc) return rc;
  int it = 0;
  double rel = resid_rel(M, n, ldm, x, b, bn, w, s), prev = rel;
  while (std::isfinite(rel) && rel > tol && it < maxiter) {
    // x is the best iterate so far; a correction that does not lower the
    // residual is rolled back
    hipMemcpyAsync(w.xbest, x, sizeof(double2) * n, hipMemcpyDeviceToDevice, s);
    convert_c128_to_c64(w.r, w.r32, n, 1, n, n, s);
    if ((rc = lu_solve<float2>(LU32, n, ld32, ipiv, 'N', 1, w.r32, n, s, &w.sws)) != 0) return rc;
    axpy_c64_into_c128(w.r32, x, n, s);
    it++;
    rel = resid_rel(M, n, ldm, x, b, bn, w, s);
    if (!(rel < prev)) {
      hipMemcpyAsync(x, w.xbest, sizeof(double2) * n, hipMemcpyDeviceToDevice, s);
      rel = prev;
      break;
    }
    prev = rel;
  }
  if (!std::isfinite(rel)) return -4;
  if (rel > tol && gmres_fallback) {
    int git = 0;
    double grel = rel;
    rc = gmres_ir(M, n, ldm, LU32, ld32, ipiv, b, x, tol, maxiter, true, &git, &grel, w, s);
    if (rc < 0) return rc;
    *iters = it + git;
    *relres = grel;
    return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return -3;
  *iters = it;
  *relres = rel;
  return rel <= tol ? 0 : 1;
}

}  // namespace scuffx
