// refine.h -- complex128 solutions from complex64 LU factors: classical
// iterative refinement (best iterate kept) and GMRES-IR.
#pragma once

#include <hip/hip_runtime.h>

#include "lu.h"

namespace scuffx {

#define REFINE_PARTS 64        // partial sums per dot product
#define REFINE_GMRES_MAX 32    // largest GMRES restart length
#define REFINE_GMRES_M 20      // restart length used by the library

// device buffers of one refinement: residual r, best iterate, GMRES
// correction d (n each), residual partial sums, complex64 vector, a
// scalar slot, dot-product partial sums, Krylov bases V (n x (m+1)) and
// Z (n x m), and the complex64 solve's buffers
struct RefineWs {
  double2 *r = nullptr, *xbest = nullptr, *d = nullptr, *work = nullptr, *part = nullptr, *V = nullptr,
          *Z = nullptr;
  float2 *r32 = nullptr;
  double *scal = nullptr;
  int m = 0;
  SolveWs<float2> sws;
};
size_t refine_ws_bytes(long n, int m);
RefineWs refine_ws_carve(char *base, long n, int m);

// Classical refinement: x0 = LU32^{-1} b, then x += LU32^{-1}(b - M x)
// while |b - M x|_inf / |b|_inf falls, at most maxiter corrections; a
// correction that raises the residual is undone.  If the tolerance is not
// reached and gmres_fallback is set, GMRES-IR continues from the best
// iterate.  Returns 0 (converged), 1 (not converged: x is the best
// iterate, *relres its residual) or a negative error (-3 device, -4 non-
// finite residual, -6 a solve's bounded wait expired).
int refine_ir(const double2 *M, long n, long ldm, const float2 *LU32, long ld32, const int *ipiv,
              const double2 *b, double2 *x, double tol, int maxiter, int gmres_fallback, int *iters,
              double *relres, RefineWs &w, hipStream_t s);

// GMRES-IR: corrections from restarted flexible GMRES(w.m) in complex128
// preconditioned by the complex64 factors; maxiter bounds the total inner
// steps.  x_given: start from x instead of LU32^{-1} b.  Same returns.
int gmres_ir(const double2 *M, long n, long ldm, const float2 *LU32, long ld32, const int *ipiv,
             const double2 *b, double2 *x, double tol, int maxiter, bool x_given, int *iters, double *relres,
             RefineWs &w, hipStream_t s);

}  // namespace scuffx
