// rhs.hip -- plane-wave right-hand side and power from the surface
// currents (the callers on either side of the system path).
//
// RHS: projection of the incident fields on each RWG function with the
// 78-point degree-20 triangle rule per panel; sources in the exterior
// region enter with a minus sign, E rows are divided by ZVAC
// (AssembleRHSVector.cc:190-246, :290-348; PlaneWave.cc:71-87).
// Power: extinction from the solution and the RHS; absorbed power from
// the crossed overlap integrals f_a.(n x f_b) of edges sharing a panel
// (OPFT.cc:82-146, :310-364).  Both reductions are two-pass and
// deterministic.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "rhs.h"

namespace scuffx {

#define ZVAC_ 376.73031346177

__constant__ double c_tcr20[234];

__device__ void pw_fields(const cplx *E0, const double *nh, cplx K, cplx Z, const double *X, cplx *EH)
{
  cplx ef = cexp_(IMUL(K) * (nh[0] * X[0] + nh[1] * X[1] + nh[2] * X[2]));
  EH[0] = E0[0] * ef; EH[1] = E0[1] * ef; EH[2] = E0[2] * ef;
  EH[3] = (nh[1] * EH[2] - nh[2] * EH[1]) / Z;
  EH[4] = (nh[2] * EH[0] - nh[0] * EH[2]) / Z;
  EH[5] = (nh[0] * EH[1] - nh[1] * EH[0]) / Z;
}

__device__ void panel_ip(const double *V1, const double *V2, const double *V3, const double *Q, double pre,
                         const cplx *E0, const double *nh, cplx K, cplx Z, double sgn, double *I)
{
  double A[3], B[3], C, J;
  for (int i = 0; i < 3; i++) { A[i] = V2[i] - V1[i]; B[i] = V3[i] - V1[i]; }
  C = A[1] * B[2] - A[2] * B[1]; J = C * C;
  C = A[2] * B[0] - A[0] * B[2]; J += C * C;
  C = A[0] * B[1] - A[1] * B[0]; J += C * C;
  J = sqrt(J);
  for (int q = 0; q < 4; q++) I[q] = 0;
  for (int q = 0; q < 78; q++) {
    double u = c_tcr20[3 * q], v = c_tcr20[3 * q + 1], w = J * c_tcr20[3 * q + 2];
    double X[3], f[3];
    for (int i = 0; i < 3; i++) X[i] = V1[i] + u * A[i] + v * B[i];
    for (int i = 0; i < 3; i++) f[i] = (X[i] - Q[i]) * pre;
    cplx EH[6];
    pw_fields(E0, nh, K, Z, X, EH);
    for (int i = 0; i < 6; i++) EH[i] = sgn * EH[i];
    cplx e = f[0] * EH[0] + f[1] * EH[1] + f[2] * EH[2];
    cplx h = f[0] * EH[3] + f[1] * EH[4] + f[2] * EH[5];
    I[0] += w * e.r; I[1] += w * e.i; I[2] += w * h.r; I[3] += w * h.i;
  }
}

// kz (optional, device): K and Z as re, im, re, im in place of the values
// passed (a recorded sequence then reads the frequency from memory)
__global__ void k_rhs(const double *V, const int *PVI, const double *PA, const int *EV, const int *EI,
                      const double *EL, int nE, int pec, double sgn, cplx e0x, cplx e0y, cplx e0z,
                      double nx, double ny, double nz, cplx K, cplx Z, const double *kz, double2 *rhs)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  if (kz) { K = mk(kz[0], kz[1]); Z = mk(kz[2], kz[3]); }
  const cplx E0[3] = {e0x, e0y, e0z};
  const double nh[3] = {nx, ny, nz};
  const double *QP = V + 3 * EV[4 * e + 2], *V1 = V + 3 * EV[4 * e], *V2 = V + 3 * EV[4 * e + 1];
  int pP = EI[4 * e], pM = EI[4 * e + 1];
  double IP[4], IM[4] = {0, 0, 0, 0};
  panel_ip(QP, V1, V2, QP, EL[e] / (2.0 * PA[pP]), E0, nh, K, Z, sgn, IP);
  if (pM >= 0) {
    const double *QM = V + 3 * EV[4 * e + 3];
    panel_ip(V1, V2, QM, QM, EL[e] / (2.0 * PA[pM]), E0, nh, K, Z, sgn, IM);
  }
  cplx EP = mk(IP[0] - IM[0], IP[1] - IM[1]), HP = mk(IP[2] - IM[2], IP[3] - IM[3]);
  if (pec) rhs[e] = make_double2(EP.r / ZVAC_, EP.i / ZVAC_);
  else { rhs[2 * e] = make_double2(EP.r / ZVAC_, EP.i / ZVAC_); rhs[2 * e + 1] = make_double2(HP.r, HP.i); }
}

static bool g_tcr20_ready[64] = {false};

#include "tcr20.inc"

hipError_t launch_rhs(hipStream_t s, const double *V, const int *PVI, const double *PA, const int *EV,
                      const int *EI, const double *EL, int nE, int pec, double sgn, const cdouble *E0,
                      const double *nhat, cdouble K, cdouble Z, double2 *rhs, const double *kz)
{
  int dev = 0;
  hipGetDevice(&dev);
  if (!g_tcr20_ready[dev & 63]) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tcr20), scuffx_tcr20, sizeof(scuffx_tcr20));
    if (e != hipSuccess) return e;
    g_tcr20_ready[dev & 63] = true;
  }
  if (nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rhs, dim3((nE + 63) / 64), dim3(64), 0, s, V, PVI, PA, EV, EI, EL, nE, pec, sgn,
                     mk(E0[0].real(), E0[0].imag()), mk(E0[1].real(), E0[1].imag()), mk(E0[2].real(), E0[2].imag()),
                     nhat[0], nhat[1], nhat[2], mk(K.real(), K.imag()), mk(Z.real(), Z.imag()), kz, rhs);
  return hipGetLastError();
}

// per-edge power contributions: out[2e] = extinction, out[2e+1] = absorbed
__global__ void k_power(const double *V, const int *PVI, const double *PA, const double *PZ, const int *PEI,
                        const int *EI, const double *EL, int nE, int pec, const double2 *KN, const double2 *RHS,
                        double *out)
{
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nE) return;
  double pext, pabs = 0.0;
  if (pec) {
    double2 k = KN[a], r = RHS[a];
    pext = 0.5 * (-ZVAC_ * (k.x * r.x + k.y * r.y));
    out[2 * a] = pext; out[2 * a + 1] = 0.0;
    return;
  }
  double2 k = KN[2 * a], r = RHS[2 * a], nn = KN[2 * a + 1], h = RHS[2 * a + 1];
  pext = 0.5 * (-ZVAC_ * (k.x * r.x + k.y * r.y) + ZVAC_ * (nn.x * h.x + nn.y * h.y));
  int pa2[2] = {EI[4 * a], EI[4 * a + 1]}, qa2[2] = {EI[4 * a + 2], EI[4 * a + 3]};
  for (int sa = 0; sa < 2; sa++) {
    int p = pa2[sa];
    if (p < 0) continue;
    for (int le = 0; le < 3; le++) {
      int b = PEI[3 * p + le];
      if (b < 0) continue;
      int sb = (EI[4 * b] == p) ? 0 : 1;
      int qb = sb == 0 ? EI[4 * b + 2] : EI[4 * b + 3];
      double sign = (sa == sb) ? 1.0 : -1.0;
      const double *Qa = V + 3 * PVI[3 * p + qa2[sa]], *Q1 = V + 3 * PVI[3 * p + (qa2[sa] + 1) % 3];
      const double *Q2 = V + 3 * PVI[3 * p + (qa2[sa] + 2) % 3], *Qb = V + 3 * PVI[3 * p + qb];
      const double *Z = PZ + 3 * p;
      double L1[3], L2[3], DQ[3], ZxDQ[3];
      for (int i = 0; i < 3; i++) { L1[i] = Q1[i] - Qa[i]; L2[i] = Q2[i] - Q1[i]; DQ[i] = Qa[i] - Qb[i]; }
      cross3(Z, DQ, ZxDQ);
      double tf = ((2.0 * L1[0] + L2[0]) * ZxDQ[0] + (2.0 * L1[1] + L2[1]) * ZxDQ[1] + (2.0 * L1[2] + L2[2]) * ZxDQ[2]) / 6.0;
      double ox = sign * EL[a] * EL[b] / (2.0 * PA[p]) * tf;
      double2 kb = KN[2 * b], nb = KN[2 * b + 1];
      // ka* nb - na* kb with n = -ZVAC * KN[odd]
      double2 ka = k, na = nn;
      double re = -ZVAC_ * ((ka.x * nb.x + ka.y * nb.y) - (na.x * kb.x + na.y * kb.y));
      pabs += 0.25 * re * ox;
    }
  }
  out[2 * a] = pext; out[2 * a + 1] = pabs;
}

__global__ void k_sum2(const double *in, int n, double *out)
{
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < n; i += 256) { a += in[2 * i]; b += in[2 * i + 1]; }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) { if (threadIdx.x < q) { s0[threadIdx.x] += s0[threadIdx.x + q]; s1[threadIdx.x] += s1[threadIdx.x + q]; } __syncthreads(); }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

hipError_t launch_power(hipStream_t s, const double *V, const int *PVI, const double *PA, const double *PZ,
                        const int *PEI, const int *EI, const double *EL, int nE, int pec, const double2 *KN,
                        const double2 *RHS, double *work, double *out2)
{
  if (nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_power, dim3((nE + 63) / 64), dim3(64), 0, s, V, PVI, PA, PZ, PEI, EI, EL, nE, pec, KN, RHS, work);
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(256), 0, s, work, nE, out2);
  return hipGetLastError();
}

// Scattered power by energy/momentum transfer (EMTPFT.cc:629-803, integral
// method EHDERIVATIVES2 :334-362) from the exterior-region matrix Mext and
// the solution x: PScat = ZVAC/2 sum_ij w_ij with, by basis-function type
// (0 = K on a PEC surface, 1 = K on a dielectric surface, 2 = N),
//   K,K: -Re(conj(x_i) x_j) Re(M_ij)     N,N: +Re(conj(x_i) x_j) Re(M_ij)
//   1,2: +Im(conj(x_i) x_j) Im(M_ij)     2,1: -Im(conj(x_i) x_j) Im(M_ij)
//   0,2 and 2,0: nothing (a PEC surface carries no N).
// One workgroup per column; partial sums per column, then one workgroup
// adds them in a fixed order (deterministic).
__global__ void __launch_bounds__(256)
k_emt_cols(const double2 *M, long ldm, const double2 *x, const signed char *bft, int n, double *part)
{
  const int j = blockIdx.x;
  const double2 xj = x[j];
  const int tj = bft[j];
  const double2 *col = M + (long)j * ldm;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int ti = bft[i];
    const double2 xi = x[i], m = col[i];
    const double re = xi.x * xj.x + xi.y * xj.y;   // Re(conj(x_i) x_j)
    const double im = xi.x * xj.y - xi.y * xj.x;   // Im(conj(x_i) x_j)
    if (ti < 2 && tj < 2) acc -= re * m.x;
    else if (ti == 2 && tj == 2) acc += re * m.x;
    else if (ti == 1 && tj == 2) acc += im * m.y;
    else if (ti == 2 && tj == 1) acc -= im * m.y;
  }
  __shared__ double s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if (threadIdx.x < q) s[threadIdx.x] += s[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[j] = s[0];
}

__global__ void k_sum1(const double *in, int n, double scale, double *out)
{
  __shared__ double s[256];
  double a = 0;
  for (int i = threadIdx.x; i < n; i += 256) a += in[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) { if (threadIdx.x < q) s[threadIdx.x] += s[threadIdx.x + q]; __syncthreads(); }
  if (threadIdx.x == 0) *out = scale * s[0];
}

hipError_t launch_emt(hipStream_t s, const double2 *Mext, long ldm, const double2 *x, const signed char *bft, int n,
                      double *work, double *out)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_emt_cols, dim3((unsigned)n), dim3(256), 0, s, Mext, ldm, x, bft, n, work);
  hipLaunchKernelGGL(k_sum1, dim3(1), dim3(256), 0, s, work, n, 0.5 * ZVAC_, out);
  return hipGetLastError();
}

}  // namespace scuffx
