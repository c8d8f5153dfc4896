// rhs.h -- plane-wave RHS and power kernels (rhs.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>

namespace scuffx {

using cdouble = std::complex<double>;

// EV[4e] = iV1, iV2, iQP, iQM ; EI[4e] = iPPanel, iMPanel, PIndex, MIndex;
// kz (optional, device): K, Z read from memory instead of the values given
hipError_t launch_rhs(hipStream_t s, const double *V, const int *PVI, const double *PA, const int *EV,
                      const int *EI, const double *EL, int nE, int pec, double sgn, const cdouble *E0,
                      const double *nhat, cdouble K, cdouble Z, double2 *rhs, const double *kz = nullptr);
// out2 = {extinction, absorbed}; work holds 2*nE doubles
hipError_t launch_power(hipStream_t s, const double *V, const int *PVI, const double *PA, const double *PZ,
                        const int *PEI, const int *EI, const double *EL, int nE, int pec, const double2 *KN,
                        const double2 *RHS, double *work, double *out2);

// EMT scattered power of the solution x from the exterior-region matrix
// Mext (n x n); bft: basis-function types (0 K on PEC, 1 K, 2 N); work: n
// doubles; *out (device) = PScat
hipError_t launch_emt(hipStream_t s, const double2 *Mext, long ldm, const double2 *x, const signed char *bft, int n,
                      double *work, double *out);

}  // namespace scuffx
