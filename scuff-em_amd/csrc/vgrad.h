// vgrad.h -- derivatives of the BEM matrix, the plane-wave RHS and the
// absorbed power with respect to the mesh vertex positions, contracted
// with caller-supplied weights inside the kernels (kern_vgrad.hip).
//
// For a weight matrix C the matrix kernels return, per vertex coordinate
// theta, the complex number sum_ij C_ij dM_ij/dtheta.  The matrix is
// symmetric, so only S = C + C^T enters.  The gradient array holds one
// complex slot per vertex coordinate: slot gofs(s) + 3 v + c for vertex v
// of surface s.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>

#include "bem.h"

namespace scuffx {

using cdouble = std::complex<double>;

// one surface as the gradient kernels see it
struct VGSurf {
  const double *V = nullptr;     // 3 per vertex
  const int *PVI = nullptr;      // 3 per panel
  const int *PBF = nullptr;      // 3 per panel: edge opposite vertex slot q, -1 if none
  const double *PSG = nullptr;   // 3 per panel: +1 positive, -1 negative panel of that edge
  const double *EL = nullptr;    // edge lengths
  int nP = 0, nV = 0, off = 0, pec = 0;   // off: first basis function of the surface
  long gofs = 0;                 // first gradient slot of the surface
};

// contraction weights: C = u v^T (C == nullptr) or a dense column-major C
struct VGW {
  const double2 *u = nullptr, *v = nullptr;
  const double2 *C = nullptr;
  long ldc = 0;
};

// a central difference of one touching or self panel pair's integrals
struct VGFDItem {
  int sa, pa, sb, pb;   // surfaces and panels (rows: sa/pa)
  int blk;              // region constants at regs[2 blk ...]
  int nreg, self;       // self: a panel with itself (weight 1/2, H1 = 0)
  int stride;           // doubles per (region, combination) in the results
  long hofs;            // results of the +h evaluation; -h follows after nreg*9*stride
  long slot;            // gradient slot of the moved coordinate
  double h;
};

// far panel pairs of rows on surface A x columns on surface B (every
// ordered pair that is not near and not a panel with itself), degree-4
// rule, derivative with respect to the row panel's vertices
hipError_t launch_vg_far(hipStream_t s, const VGSurf &A, const VGSurf &B, int same, const int *nearRowPtr,
                         const int *nearCol, int nreg, const RegionK *regs, int needH1, const VGW &w,
                         double2 *grad);
// listed ordered near pairs without a common vertex (pairs[2 i], pairs[2 i + 1]),
// 78-point rule on each quarter of both panels (312 x 312 point pairs)
hipError_t launch_vg_nearsep(hipStream_t s, const VGSurf &A, const VGSurf &B, const int *pairs, int npairs,
                             int nreg, const RegionK *regs, int needH1, const VGW &w, double2 *grad);
// central differences of touching / self pairs from the results of the
// Taylor-Duffy kernels (HN: touching, HS: self)
hipError_t launch_vg_fd(hipStream_t s, const VGSurf *surfs, const VGFDItem *items, int nitems,
                        const RegionK *regs, const double *HN, const double *HS, const VGW &w, double2 *grad);
// rho_i = sum_j (C_ij M_ij + C_ji M_ji) into rho (zeroed by the caller)
hipError_t launch_vg_rho(hipStream_t s, const double2 *M, long ldm, int n, const VGW &w, double2 *rho);
// edge-length term: sum over the rows of each edge of rho, times d ln L / dtheta
hipError_t launch_vg_len(hipStream_t s, const VGSurf &A, const int *EV, int nE, const double2 *rho,
                         double2 *grad);
// sum_i mu_i d b_i / dtheta for the plane-wave RHS of one surface
// (sgn, E0, nhat, K, Z as for the RHS kernel)
hipError_t launch_vg_rhs(hipStream_t s, const VGSurf &A, const int *EV, int nE, double sgn, const cdouble *E0,
                         const double *nhat, cdouble K, cdouble Z, const double2 *mu, double2 *grad);
// absorbed power (crossed overlaps) of the currents KN on one non-PEC
// surface: its explicit derivative into gradR (real, 3 per vertex) and
// gamma with dPAbs = Re(gamma^T dKN) (gamma zeroed by the caller)
hipError_t launch_vg_pabs(hipStream_t s, const VGSurf &A, const double *PZ, const int *PEI, const int *EI, int nE,
                          const double2 *KN, double *gradR, double2 *gamma);

}  // namespace scuffx
