// adjoint.h -- derivative-block contractions for adjoint gradients
#pragma once

#include <hip/hip_runtime.h>

namespace scuffx {

// partial[k], k < nparts: workgroup k's share of
// sum_ij G_ij (conj(xa_i) yb_j + conj(xb_j) ya_i) over the na x nb block G
hipError_t launch_adj_contract(hipStream_t s, const double2 *G, long ldg, int na, int nb, const double2 *xa,
                               const double2 *ya, const double2 *xb, const double2 *yb, double2 *partial,
                               int nparts);

}  // namespace scuffx
