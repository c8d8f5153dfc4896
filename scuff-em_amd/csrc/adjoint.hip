// adjoint.hip -- contraction of the derivative blocks of the BEM matrix with
// two vectors, for the adjoint gradient of a solve with respect to a rigid
// motion of one surface.
//
// For a motion of surface a, only the blocks (a, b) and (b, a), b != a,
// change; the matrix is symmetric, so the (b, a) block's derivative is the
// transpose of the (a, b) one.  Per derivative block G (na x nb, column-
// major) this sums
//     sum_ij G_ij (conj(x_a,i) y_b,j + conj(x_b,j) y_a,i)
// as per-workgroup partial sums in a fixed order (deterministic); the
// host adds the partials.
#include <hip/hip_runtime.h>

#include "adjoint.h"

namespace scuffx {

__global__ void __launch_bounds__(256)
k_adj_contract(const double2 *G, long ldg, int na, int nb, const double2 *xa, const double2 *ya,
               const double2 *xb, const double2 *yb, double2 *partial)
{
  __shared__ double sr[256], si[256];
  const long total = (long)na * nb;
  double ar = 0, ai = 0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % na), j = (int)(e / na);
    const double2 g = G[i + (long)j * ldg];
    const double2 p = xa[i], q = yb[j], u = xb[j], v = ya[i];
    // w = conj(p) q + conj(u) v
    const double wr = p.x * q.x + p.y * q.y + u.x * v.x + u.y * v.y;
    const double wi = p.x * q.y - p.y * q.x + u.x * v.y - u.y * v.x;
    ar += g.x * wr - g.y * wi;
    ai += g.x * wi + g.y * wr;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sr[0], si[0]);
}

hipError_t launch_adj_contract(hipStream_t s, const double2 *G, long ldg, int na, int nb, const double2 *xa,
                               const double2 *ya, const double2 *xb, const double2 *yb, double2 *partial,
                               int nparts)
{
  if (na <= 0 || nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_contract, dim3((unsigned)nparts), dim3(256), 0, s, G, ldg, na, nb, xa, ya, xb, yb,
                     partial);
  return hipGetLastError();
}

}  // namespace scuffx
