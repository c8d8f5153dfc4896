// rhs.h -- plane-wave RHS and power kernels (rhs.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>

namespace scuffx {

using cdouble = std::complex<double>;

// EV[4e] = iV1, iV2, iQP, iQM ; EI[4e] = iPPanel, iMPanel, PIndex, MIndex
hipError_t launch_rhs(hipStream_t s, const double *V, const int *PVI, const double *PA, const int *EV,
                      const int *EI, const double *EL, int nE, int pec, double sgn, const cdouble *E0,
                      const double *nhat, cdouble K, cdouble Z, double2 *rhs);
// out2 = {extinction, absorbed}; work holds 2*nE doubles
hipError_t launch_power(hipStream_t s, const double *V, const int *PVI, const double *PA, const double *PZ,
                        const int *PEI, const int *EI, const double *EL, int nE, int pec, const double2 *KN,
                        const double2 *RHS, double *work, double *out2);

}  // namespace scuffx
