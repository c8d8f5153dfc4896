// kern_self.hip -- self (common-triangle) panel integrals by the
// generalized Taylor-Duffy method, every frequency.
//
// For a panel paired with itself the reference calls TaylorDuffy with
// P = {1, (x-Q).(x'-Q')}, the Helmholtz kernel e^{ikr}/4 pi r, and
// H0 = I_1 - 4 I_0/k^2, H1 = 0 (PanelPanelInteractions.cc:449-491).  The
// remaining 1-D integral is done by p-adaptive Clenshaw-Curtis
// quadrature, relative tolerance 1e-10, at most 1000 new points
// (TaylorDuffy.cc:152-367, libs/libSGJC/pcubature.c:323-385).
// One wave per (panel, Qa, Qb, region) job; the function values of a job
// live in a scratch slab so every refinement level reuses them.
#include <hip/hip_runtime.h>

#include "bem.h"

namespace scuffx {

#define TD_NWP 5
#define TD_MAXM 9            // 2^(9+1)+1 = 1025 nodes reached by MaxEval=1000
#define TD_NODES 1025
#define TD_NPRE 64           // cache entries computed up front, one per lane

__constant__ double c_ccx[512];            // nodes x_i, layout as the oracle
__constant__ double c_ccw[TD_MAXM + 1 + (1 << (TD_MAXM + 1))];

// relative exponential ExpRelV3P0 (TaylorDuffy.cc:840-893)
__device__ cplx exprel_v3(int n, cplx Z)
{
  int m;
  if (Z.i == 0.0) {
    double X = Z.r, Term, Sum;
    if (fabs(X) < 0.1) {
      Sum = 1.0;
      for (Term = 1.0, m = 1; m < 100; m++) {
        Term *= X / ((double)(m + n));
        Sum += Term;
        if (fabs(Term) < 1.0e-8 * fabs(Sum)) break;
      }
      return mk(Sum, 0);
    }
    Sum = exp(X);
    for (Term = 1.0, m = 0; m < n; m++) { Sum -= Term; Term *= X / ((double)(m + 1)); }
    return mk(Sum / Term, 0);
  }
  cplx Term, Sum;
  if (sqrt(Z.r * Z.r + Z.i * Z.i) < 0.1) {
    Sum = Term = mk(1, 0);
    for (m = 1; m < 100; m++) {
      Term = Term * (Z / (double)(m + n));
      Sum += Term;
      if (norm2(Term) < 1.0e-16 * norm2(Sum)) break;
    }
    return Sum;
  }
  Sum = cexp_(Z);
  Term = mk(1, 0);
  for (m = 0; m < n; m++) { Sum -= Term; Term = Term * (Z / (double)(m + 1)); }
  return Sum / Term;
}

// common-triangle Upsilon coefficients (TaylorDuffy.cc:966-1057):
// U[d][n][0..2] = coefficients of 1, y, y^2; only M13,M14,M23,M24,V,S
// are non-zero for P in {1, PMCHWG1}.
struct CTPoly { double U[3][TD_NWP][3]; int nMin, nMax; };

__device__ void ct_upsilon(double M13, double M14, double M23, double M24, double V1, double V2,
                           double V3, double V4, double S, CTPoly &P)
{
  const double M11 = 0, M12 = 0, M22 = 0, M33 = 0, M34 = 0, M44 = 0;
  for (int d = 0; d < 3; d++) for (int n = 0; n < TD_NWP; n++) for (int j = 0; j < 3; j++) P.U[d][n][j] = 0.0;
  double (*U)[TD_NWP][3] = P.U;
  U[0][0][0] = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
               + M34/4 + M44/12 + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3 + S;
  U[0][1][1] = M12/3 + M14/3 + M22/6 + M23/3 + M24/3 + M34/3 + M44/6 + V2/2 + V4/2;
  U[0][1][0] = -(2*M11)/3 - (5*M12)/6 - (4*M13)/3 - (5*M14)/6 - M22/3 - (5*M23)/6
               - (2*M24)/3 - (2*M33)/3 - (5*M34)/6 - M44/3 - (3*V1)/2 - V2 - (3*V3)/2 - V4 - 2*S;
  U[0][2][2] = M22/4 + M44/4;
  U[0][2][1] = -M12/2 - M14 - M22/2 - M23 - M24 - M34/2 - M44/2 - V2 - V4;
  U[0][2][0] = (3*M11)/4 + M12 + M13 + M14 + M22/2 + M23 + M24 + (3*M33)/4 + M34 + M44/2
               + V1 + V2 + V3 + V4 + S;
  U[0][3][2] = -M22/2 - M44/2;
  U[0][3][1] = M14 + M22/2 + M23 + M24 + M44/2 + V2/2 + V4/2;
  U[0][3][0] = -M11/2 - M12/2 - M14/2 - M22/3 - M23/2 - (2*M24)/3 - M33/2 - M34/2 - M44/3
               - V1/6 - V2/3 - V3/6 - V4/3;
  U[0][4][2] = M22/4 + M44/4;
  U[0][4][1] = M12/6 - M14/3 - M22/6 - M23/3 - M24/3 + M34/6 - M44/6;
  U[0][4][0] = M11/6 + M12/12 - M13/6 + M14/12 + M22/12 + M23/12 + M24/6 + M33/6 + M34/12 + M44/12;

  U[1][0][0] = U[0][0][0];
  U[1][1][1] = -M11/3 - M12/2 - (2*M13)/3 - M14/2 - M22/6 - M23/2 - M24/3 - M33/3
               - M34/2 - M44/6 - V1/2 - V2/2 - V3/2 - V4/2;
  U[1][1][0] = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3 - M33/3
               - M34/3 - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
  U[1][2][2] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
  U[1][2][1] = M11/2 + M12/2 + M13 + M14 + M23 + M24 + M33/2 + M34/2 + V1 + V2 + V3 + V4;
  U[1][2][0] = M22/4 + M44/4 + S;
  U[1][3][2] = -M11/2 - M12 - M22/2 - M33/2 - M34 - M44/2;
  U[1][3][1] = M12/2 - M14/2 + M22/2 - M23/2 - M24 + M34/2 + M44/2 - V1/2 - V2/2 - V3/2 - V4/2;
  U[1][3][0] = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
  U[1][4][2] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
  U[1][4][1] = -M11/6 - M12/2 - M13/3 - M22/3 + M24/3 - M33/6 - M34/2 - M44/3;
  U[1][4][0] = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6 + M33/12 + M34/12 + M44/6;

  U[2][0][0] = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
               + M34/4 + M44/12 + S + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3;
  U[2][1][1] = -M11/3 - M12/6 - (2*M13)/3 - M14/6 - M23/6 - M33/3 - M34/6 - V1/2 - V3/2;
  U[2][1][0] = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3 - M33/3 - M34/3
               - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
  U[2][2][2] = M11/4 + M33/4;
  U[2][2][1] = M11/2 + M12/2 + M13 + M33/2 + M34/2 + V1 + V3;
  U[2][2][0] = M22/4 + M44/4 + S;
  U[2][3][2] = -M11/2 - M33/2;
  U[2][3][1] = -M12/2 + M14/2 + M23/2 - M34/2 - V1/2 - V3/2;
  U[2][3][0] = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
  U[2][4][2] = M11/4 + M33/4;
  U[2][4][1] = -M11/6 + M12/6 - M13/3 - M14/3 - M23/3 - M33/6 + M34/6;
  U[2][4][0] = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6 + M33/12 + M34/12 + M44/6;

  // range of w-powers with a coefficient above 1e-8 (TaylorDuffy.cc:1507-1520)
  int nMin = TD_NWP - 1, nMax = -1;
  for (int n = 0; n < TD_NWP; n++)
    for (int d = 0; d < 3; d++)
      for (int j = 0; j < 3; j++)
        if (fabs(U[d][n][j]) > 1.0e-8) { if (n < nMin) nMin = n; if (n > nMax) nMax = n; }
  P.nMin = nMin; P.nMax = nMax;
}

struct CTJob { double A2, AdB, B2; CTPoly P[2]; cplx k; };

// integrand at y: two complex values (P=1, P=PMCHWG1) as 4 doubles.  The
// job (J, in LDS) is shared by the wave; the kernel values K_n(X_d) =
// e^{ikX} ExpRel_n(-ikX) / (n X) are formed where the sum uses them (the
// same operations as tabulating them first, without an array indexed by
// the run-time power range, which would live in scratch memory), and
// e^{ikX_d} once for both polynomials.
__device__ void ct_integrand(const CTJob &J, double y, double *f)
{
  double X[3], u1, u2;
  u1 = 1.0; u2 = y;       X[0] = sqrt(J.A2 * u1 * u1 + 2.0 * J.AdB * u1 * u2 + J.B2 * u2 * u2);
  u1 = y;   u2 = (y - 1); X[1] = sqrt(J.A2 * u1 * u1 + 2.0 * J.AdB * u1 * u2 + J.B2 * u2 * u2);
  u1 = y;   u2 = 1.0;     X[2] = sqrt(J.A2 * u1 * u1 + 2.0 * J.AdB * u1 * u2 + J.B2 * u2 * u2);
  const cplx IK = IMUL(J.k);
  cplx IKX[3], E[3];
  for (int d = 0; d < 3; d++) {
    IKX[d] = IK * X[d];
    E[d] = cexp_(IKX[d]);
  }
  for (int pk = 0; pk < 2; pk++) {
    const CTPoly &P = J.P[pk];
    cplx s = mk(0, 0);
    for (int n = P.nMin; n <= P.nMax; n++)
      for (int d = 0; d < 3; d++) {
        const cplx K = E[d] * exprel_v3(n + 1, -IKX[d]) / ((n + 1) * X[d]);
        s += (P.U[d][n][0] + y * P.U[d][n][1]) * K;
      }
    s = s * (1.0 / (4.0 * M_PI));
    f[2 * pk] = s.r; f[2 * pk + 1] = s.i;
  }
}

// weight of cache entry e in the Clenshaw-Curtis rule of the given order
// (entries: center, the pair at x0, then level l >= 1 pairs from 2^l + 1),
// 0 for entries of levels above the order; order < 0: one-point rule
__device__ inline double cc_weight(int e, int order)
{
  if (order < 0) return e == 0 ? 2.0 : 0.0;
  const double *w = c_ccw + order + (1 << order) - 1;
  if (e == 0) return w[0];
  if (e < 3) return w[1];
  const int l = 31 - __clz(e - 1);          // level of the entry
  if (l > order) return 0.0;
  const int nl = 1 << (l - 1);
  return w[1 + nl + ((e - (1 << l) - 1) >> 1)];
}

// One wave per (panel, Qa, Qb, region) job: each refinement level's new
// points are evaluated across the 64 lanes, and the two rule estimates
// are lane-strided partial sums reduced over the wave (the same stopping
// rule as the reference's p-adaptive cubature: relative 1e-10 per
// component, at most 1000 new points).
__global__ void __launch_bounds__(64)
k_self(const double *V, const int *PVI, int nP, int nreg, const RegionK *regs,
       double *HS, double *scratch, int job0, int njobs)
{
  const int t = blockIdx.x, lane = threadIdx.x;
  if (t >= njobs) return;
  const int job = job0 + t;
  const int r = job % nreg, q = (job / nreg) % 9, p = job / (9 * nreg);
  const int qa = q / 3, qb = q % 3;
  const double *V1 = V + 3 * PVI[3 * p], *V2 = V + 3 * PVI[3 * p + 1], *V3 = V + 3 * PVI[3 * p + 2];
  const double *Q = V + 3 * PVI[3 * p + qa], *QP = V + 3 * PVI[3 * p + qb];
  double A[3], B[3], D[3], DP[3];
  sub3(V2, V1, A); sub3(V3, V2, B); sub3(V1, Q, D); sub3(V1, QP, DP);
  // the job's constants in LDS (read by every lane at the same address)
  __shared__ CTJob J;
  if (lane == 0) {
    J.A2 = dot3(A, A); J.AdB = dot3(A, B); J.B2 = dot3(B, B);
    J.k = regs[r].k;
    ct_upsilon(0, 0, 0, 0, 0, 0, 0, 0, 1.0, J.P[0]);
    ct_upsilon(dot3(A, A), dot3(A, B), dot3(B, A), dot3(B, B), dot3(A, DP), dot3(B, DP),
               dot3(D, A), dot3(D, B), dot3(D, DP), J.P[1]);
  }
  __syncthreads();
  double *vals = scratch + (long long)t * TD_NODES * 4;
  const double c = 0.5, rr = 0.5, Vol = 0.5;
  // The first TD_NPRE cache entries (centre, the pair at x_0, then
  // 2^(l-1) pairs per level l: rule orders 0 .. 4 and all but the last
  // entry of order 5) are filled at once, one per lane, before the first
  // convergence test: the lanes a low order leaves idle evaluate the next
  // orders' points meanwhile.  The tests below read the same values in
  // the same order as if each order's points were computed after the
  // previous order's test; only the number of sequential passes shrinks.
  for (int e = lane; e < TD_NPRE; e += 64) {
    double y;
    if (e == 0) y = c;
    else if (e < 3) y = e == 1 ? c + rr * c_ccx[0] : c - rr * c_ccx[0];
    else {
      const int l = 31 - __clz(e - 1), k = e - (1 << l) - 1, nl = 1 << (l - 1);
      const double x = c_ccx[nl + (k >> 1)];
      y = (k & 1) ? c - rr * x : c + rr * x;
    }
    ct_integrand(J, y, vals + 4 * e);
  }
  __syncthreads();
  int m = 0, nstored = 3;
  long numEval = 0;
  double val[4], val1[4];
  for (;;) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    for (int e = lane; e < nstored; e += 64) {
      const double wa = cc_weight(e, m) * Vol, wb = cc_weight(e, m - 1) * Vol;
      for (int cc = 0; cc < 4; cc++) { a[cc] += vals[4 * e + cc] * wa; b[cc] += vals[4 * e + cc] * wb; }
    }
    for (int off = 32; off > 0; off >>= 1)
      for (int cc = 0; cc < 4; cc++) { a[cc] += __shfl_xor(a[cc], off); b[cc] += __shfl_xor(b[cc], off); }
    bool conv = true;
    for (int cc = 0; cc < 4; cc++) {
      val[cc] = a[cc]; val1[cc] = b[cc];
      const double e = fabs(val[cc] - val1[cc]);
      if (e > 0.0 && e > fabs(val[cc]) * 1.0e-10) conv = false;
    }
    if (conv || numEval > 1000 || m >= TD_MAXM) break;
    m += 1;
    const int nl = 1 << (m - 1);
    for (int k = lane; k < 2 * nl; k += 64) {
      if (nstored + k < TD_NPRE) continue;   // computed up front
      const double x = c_ccx[nl + (k >> 1)];
      ct_integrand(J, (k & 1) ? c - rr * x : c + rr * x, vals + 4 * (nstored + k));
    }
    nstored += 2 * nl;
    numEval += 2 * nl;
    __syncthreads();
  }
  if (lane == 0) {
    cplx I0 = mk(val[0], val[1]), I1 = mk(val[2], val[3]);
    cplx k = J.k;
    cplx H0 = I1 - 4.0 * I0 / (k * k);
    double *o = HS + ((long long)p * nreg + r) * 18 + 2 * q;
    o[0] = H0.r; o[1] = H0.i;
  }
}

static bool g_cc_ready[64] = {false};

hipError_t launch_self(hipStream_t stream, const double *V, const int *PVI, int nP, int nreg,
                       const RegionK *regs, double *HS, double *scratch, int chunk)
{
  int dev = 0;
  hipGetDevice(&dev);
  if (!g_cc_ready[dev & 63]) {
    // Clenshaw-Curtis nodes and weights in the nested layout of
    // pcubature.c: x[0]=1; level m>=1 at offset 2^(m-1); weights of
    // rule order m (2^(m+1) intervals) at offset m+2^m-1:
    // [center, level-0 node, level-1 nodes, ...]
    static double hx[512], hw[TD_MAXM + 1 + (1 << (TD_MAXM + 1))];
    const long double pi = 3.141592653589793238462643383279502884L;
    hx[0] = 1.0;
    for (int m = 1; m <= TD_MAXM; m++)
      for (int i = 0; i < (1 << (m - 1)); i++)
        hx[(1 << (m - 1)) + i] = (double)cosl((2 * i + 1) * pi / (long double)(1L << (m + 1)));
    auto wt = [&](long n, long j) {
      long double s = 0;
      for (long k = 1; k <= n / 2; k++) {
        long double b = (k == n / 2) ? 1.0L : 2.0L;
        s += b / (4.0L * k * k - 1.0L) * cosl(2.0L * pi * (long double)((j * k) % n) / (long double)n);
      }
      long double cj = (j == 0 || j == n) ? 1.0L : 2.0L;
      return (double)(cj / (long double)n * (1.0L - s));
    };
    for (int m = 0; m <= TD_MAXM; m++) {
      long n = 1L << (m + 1);
      double *w = hw + m + (1 << m) - 1;
      w[0] = wt(n, n / 2);
      w[1] = wt(n, 0);
      int q = 2;
      for (int l = 1; l <= m; l++)
        for (int i = 0; i < (1 << (l - 1)); i++) w[q++] = wt(n, (2 * i + 1) * (n >> (l + 1)));
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ccx), hx, sizeof(hx));
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_ccw), hw, sizeof(hw));
    if (e != hipSuccess) return e;
    g_cc_ready[dev & 63] = true;
  }
  int njobs = nP * 9 * nreg;
  for (int j0 = 0; j0 < njobs; j0 += chunk) {
    int nj = njobs - j0 < chunk ? njobs - j0 : chunk;
    hipLaunchKernelGGL(k_self, dim3(nj), dim3(64), 0, stream, V, PVI, nP, nreg, regs, HS, scratch, j0, nj);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace scuffx
