// kern_pairs.hip -- panel-pair integrals for far and near pairs, the
// edge-pair combination and stamping into the BEM matrix, and the
// symmetric mirror of the upper triangle.
//
// Far pairs (relative centroid distance > 4) use the 6x6-point fixed
// rule of GetPPIs_Cubature (PanelPanelInteractions.cc:211-350, reached at
// :420-424).  Instead of re-running the 36-point sum for each of the 9
// (Qa,Qb) source-vertex choices, one pass accumulates moments in panel-
// local coordinates xi = x - c_a, xi' = x' - c_b (R = x - x'):
//   H0(Qa,Qb) = S_xx' - Qb.S_x - Qa.S_x' + (Qa.Qb + 4/(ik)^2) S_1
//   H1(Qa,Qb) = s - Qb.u - Qa.v + (Qa x Qb).w
// with S_1 = sum Phi, S_x = sum xi Phi, S_x' = sum xi' Phi,
// S_xx' = sum (xi.xi') Phi, s = sum ((xi x xi').R) Psi,
// u = sum (R x xi) Psi, v = sum (xi' x R) Psi, w = sum R Psi, and
// Qa, Qb measured from the same panel centroids.  This is an algebraic
// rearrangement of (x-Qa).(x'-Qb) and ((x-Qa)x(x'-Qb)).R; the 9 results
// equal the reference's per-Q sums up to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bem.h"

namespace scuffx {

__constant__ double c_tcr4[18] = SCUFFX_TCR4_INIT;

#define SCUFFX_INV4PI 0.079577471545947667884   // 1/(4 pi)
// first word of a band record whose pair is near (its values live in the
// near buffer): a NaN bit pattern no integral produces
#define SCUFFX_NEAR_MARK 0x7FF4DEADBEEF0001ULL

// binary search of sorted nearCol[lo..hi) for pb; index or -1
__device__ inline int near_find(const int *rowPtr, const int *col, int pa, int pb)
{
  int lo = rowPtr[pa], hi = rowPtr[pa + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    int c = col[mid];
    if (c == pb) return mid;
    if (c < pb) lo = mid + 1; else hi = mid;
  }
  return -1;
}

// moments of one panel pair for one wavenumber (far: Phi = e^{ikr}/4pi r)
struct Moments {
  cplx S1, Sx[3], Sy[3], Sxy;
  cplx s, u[3], v[3], w[3];
};

template <int DESING>
__device__ inline void pair_moments(const double *xi, const double *eta, const double *D,
                                    cplx k, int needH1, Moments &m)
{
#pragma clang fp contract(fast)
  m.S1 = mk(0, 0); m.Sxy = mk(0, 0); m.s = mk(0, 0);
  for (int c = 0; c < 3; c++) { m.Sx[c] = mk(0, 0); m.Sy[c] = mk(0, 0); m.u[c] = mk(0, 0); m.v[c] = mk(0, 0); m.w[c] = mk(0, 0); }
  cplx ik = IMUL(k);
  for (int i = 0; i < 6; i++) {
    const double wi = c_tcr4[3 * i + 2];
    const double *x = xi + 3 * i;
    for (int j = 0; j < 6; j++) {
      const double wj = c_tcr4[3 * j + 2];
      const double *y = eta + 3 * j;
      double R[3] = {D[0] + x[0] - y[0], D[1] + x[1] - y[1], D[2] + x[2] - y[2]};
      double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      double r = sqrt(r2);
      double ir = 1.0 / r;
      cplx Phi;
      if (DESING) {
        // ExpRel(ikr,4) = sum_{m>=4} (ikr)^m/m!, stopped at |term|<1e-8|sum|
        // (PanelPanelInteractions.cc:84-103)
        cplx z = ik * r, term = mk(1, 0), sum = mk(0, 0);
        int q;
        for (q = 1; q < 4; q++) term = term * (z / (double)q);
        for (; q < 100; q++) {
          term = term * (z / (double)q);
          sum += term;
          if (norm2(term) < 1.0e-16 * norm2(sum)) break;
        }
        Phi = sum * (SCUFFX_INV4PI * ir);
      } else {
        double e = exp(-k.i * r), sn, cs;
        sincos(k.r * r, &sn, &cs);
        Phi = mk(e * cs, e * sn) * (SCUFFX_INV4PI * ir);
      }
      Phi = Phi * (wi * wj);
      m.S1 += Phi;
      for (int c = 0; c < 3; c++) { m.Sx[c] += x[c] * Phi; m.Sy[c] += y[c] * Phi; }
      m.Sxy += (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) * Phi;
      if (needH1) {
        cplx Psi = Phi * (ik - mk(ir, 0)) * ir;
        double xy[3], Rx[3], yR[3];
        cross3(x, y, xy);
        cross3(R, x, Rx);
        cross3(y, R, yR);
        m.s += (xy[0] * R[0] + xy[1] * R[1] + xy[2] * R[2]) * Psi;
        for (int c = 0; c < 3; c++) { m.u[c] += Rx[c] * Psi; m.v[c] += yR[c] * Psi; m.w[c] += R[c] * Psi; }
      }
    }
  }
}

// one (Qa, Qb) source-vertex combination of a pair's moments
__device__ inline void combo_ab(const Moments &m, const double *qa, const double *qb, cplx c4, int needH1, cplx *h0,
                                cplx *h1)
{
  cplx h = m.Sxy;
  for (int c = 0; c < 3; c++) h = h - qb[c] * m.Sx[c] - qa[c] * m.Sy[c];
  h += (mk(qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2], 0) + c4) * m.S1;
  *h0 = h;
  if (needH1) {
    double qq[3];
    cross3(qa, qb, qq);
    cplx t = m.s;
    for (int c = 0; c < 3; c++) t = t - qb[c] * m.u[c] - qa[c] * m.v[c] + qq[c] * m.w[c];
    *h1 = t;
  }
}

__device__ inline void combos_from_moments(const Moments &m, const double *Qa, const double *Qb,
                                           cplx k, int needH1, cplx *H0, cplx *H1)
{
  cplx ik = IMUL(k);
  cplx c4 = 4.0 / (ik * ik);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) combo_ab(m, Qa + 3 * a, Qb + 3 * b, c4, needH1, H0 + 3 * a + b, H1 + 3 * a + b);
}

// TCR4 point i of a panel given as p = {V0 - centroid, V1 - V0, V2 - V0}
// (3 each), panel-local: p0 + s pA + u pB, rounded term by term
__device__ inline void tcr4_point(const double *p, int i, double *x)
{
  const double s = c_tcr4[3 * i], u = c_tcr4[3 * i + 1];
  for (int c = 0; c < 3; c++) x[c] = p[c] + s * p[3 + c] + u * p[6 + c];
}

// pair_moments<0> with the points formed on the fly and the 6 x 6 point
// loop kept rolled: the same sums in the same order, in about half the
// registers of the unrolled form (two waves per SIMD instead of one)
__device__ inline void pair_moments_far(const double *pa, const double *pb, const double *D, cplx k, int needH1,
                                        Moments &m)
{
#pragma clang fp contract(fast)
  m.S1 = mk(0, 0); m.Sxy = mk(0, 0); m.s = mk(0, 0);
  for (int c = 0; c < 3; c++) { m.Sx[c] = mk(0, 0); m.Sy[c] = mk(0, 0); m.u[c] = mk(0, 0); m.v[c] = mk(0, 0); m.w[c] = mk(0, 0); }
  cplx ik = IMUL(k);
#pragma unroll 1
  for (int i = 0; i < 6; i++) {
    const double wi = c_tcr4[3 * i + 2];
    double x[3];
    tcr4_point(pa, i, x);
#pragma unroll 1
    for (int j = 0; j < 6; j++) {
      const double wj = c_tcr4[3 * j + 2];
      double y[3];
      tcr4_point(pb, j, y);
      double R[3] = {D[0] + x[0] - y[0], D[1] + x[1] - y[1], D[2] + x[2] - y[2]};
      double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      double r = sqrt(r2);
      double ir = 1.0 / r;
      // (a real wavenumber, as in a lossless region at real frequency,
      // skips the exponential: exp(-0) is 1 exactly)
      double e = 1.0, sn, cs;
      if (k.i != 0) e = exp(-k.i * r);
      sincos(k.r * r, &sn, &cs);
      cplx Phi = mk(e * cs, e * sn) * (SCUFFX_INV4PI * ir);
      Phi = Phi * (wi * wj);
      m.S1 += Phi;
      for (int c = 0; c < 3; c++) { m.Sx[c] += x[c] * Phi; m.Sy[c] += y[c] * Phi; }
      m.Sxy += (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) * Phi;
      if (needH1) {
        cplx Psi = Phi * (ik - mk(ir, 0)) * ir;
        double xy[3], Rx[3], yR[3];
        cross3(x, y, xy);
        cross3(R, x, Rx);
        cross3(y, R, yR);
        m.s += (xy[0] * R[0] + xy[1] * R[1] + xy[2] * R[2]) * Psi;
        for (int c = 0; c < 3; c++) { m.u[c] += Rx[c] * Psi; m.v[c] += yR[c] * Psi; m.w[c] += R[c] * Psi; }
      }
    }
  }
}

// ---------------------------------------------------------------------
// K1: far panel pairs of one column band.  The band is a list of column
// panels (slots); every row panel pa < rowLimit is paired with every slot
// panel pb in the given order (row panel first), skipping the pair of a
// panel with itself and the near pairs.  Record of (pa, slot) at
// HB[(slot * rowsCap + pa) * nreg * 9 * stride]: for each region the 9
// (Qa, Qb) combinations 3 qa + qb of {H0[, H1]}.
// grid.x over row panels, grid.y = slot.
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_far(const double *Va, const int *PVIa, const double *PCa, int rowLimit,
      const double *Vb, const int *PVIb, const double *PCb, const int *slotPanels, int same,
      const int *nearRowPtr, const int *nearCol, int nreg, const RegionK *regs,
      int needH1, double *HB, long rowsCap)
{
  const int slot = blockIdx.y;
  const int pa = blockIdx.x * blockDim.x + threadIdx.x;
  if (pa >= rowLimit) return;
  const int pb = slotPanels[slot];
  if (same && pb == pa) return;
  if (near_find(nearRowPtr, nearCol, pa, pb) >= 0) {
    const int stride = needH1 ? 4 : 2;
    HB[((long)slot * rowsCap + pa) * (long)(nreg * 9 * stride)] = __longlong_as_double((long long)SCUFFX_NEAR_MARK);
    return;
  }
  double ca[3] = {PCa[4 * pa], PCa[4 * pa + 1], PCa[4 * pa + 2]};
  double cb[3] = {PCb[4 * pb], PCb[4 * pb + 1], PCb[4 * pb + 2]};
  const double *A0 = Va + 3 * PVIa[3 * pa], *A1 = Va + 3 * PVIa[3 * pa + 1], *A2 = Va + 3 * PVIa[3 * pa + 2];
  const double *B0 = Vb + 3 * PVIb[3 * pb], *B1 = Vb + 3 * PVIb[3 * pb + 1], *B2 = Vb + 3 * PVIb[3 * pb + 2];
  // panels as {V0 - centroid, V1 - V0, V2 - V0}; D = centroid difference
  double PA[9], PB[9], D[3];
  for (int c = 0; c < 3; c++) {
    PA[c] = A0[c] - ca[c]; PA[3 + c] = A1[c] - A0[c]; PA[6 + c] = A2[c] - A0[c];
    PB[c] = B0[c] - cb[c]; PB[3 + c] = B1[c] - B0[c]; PB[6 + c] = B2[c] - B0[c];
    D[c] = ca[c] - cb[c];
  }
  const int stride = needH1 ? 4 : 2;
  double *out = HB + ((long)slot * rowsCap + pa) * (long)(nreg * 9 * stride);
  for (int r = 0; r < nreg; r++) {
    Moments m;
    cplx k = regs[r].k;
    pair_moments_far(PA, PB, D, k, needH1, m);
    // source vertices relative to the centroids
    double Qa[9], Qb[9];
    for (int c = 0; c < 3; c++) {
      Qa[c] = A0[c] - ca[c]; Qa[3 + c] = A1[c] - ca[c]; Qa[6 + c] = A2[c] - ca[c];
      Qb[c] = B0[c] - cb[c]; Qb[3 + c] = B1[c] - cb[c]; Qb[6 + c] = B2[c] - cb[c];
    }
    cplx ik = IMUL(k);
    cplx c4 = 4.0 / (ik * ik);
    for (int q = 0; q < 9; q++) {
      cplx h0, h1;
      combo_ab(m, Qa + 3 * (q / 3), Qb + 3 * (q % 3), c4, needH1, &h0, &h1);
      double *o = out + (r * 9 + q) * stride;
      o[0] = h0.r; o[1] = h0.i;
      if (needH1) { o[2] = h1.r; o[3] = h1.i; }
    }
  }
}

// ---------------------------------------------------------------------
// K1 in single precision (the mixed form's optional far pairs): the same
// records from the same 36-point moments, with the point arithmetic, the
// kernel values and the sums in FP32 (panel-local coordinates, so the
// lost digits are relative to the panel size and the centroid distance).
// Each far entry then carries ~1e-7 relative error instead of ~1e-16.
// ---------------------------------------------------------------------
struct fcx {
  float r, i;
};
__device__ inline fcx fmk(float r, float i) { fcx z; z.r = r; z.i = i; return z; }
__device__ inline fcx operator+(fcx a, fcx b) { return fmk(a.r + b.r, a.i + b.i); }
__device__ inline fcx operator-(fcx a, fcx b) { return fmk(a.r - b.r, a.i - b.i); }
__device__ inline fcx operator*(fcx a, fcx b) { return fmk(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__device__ inline fcx operator*(float s, fcx a) { return fmk(s * a.r, s * a.i); }
__device__ inline fcx &operator+=(fcx &a, fcx b) { a.r += b.r; a.i += b.i; return a; }

struct Moments32 {
  fcx S1, Sx[3], Sy[3], Sxy;
  fcx s, u[3], v[3], w[3];
};

// the far-pair moments in FP32 from panels given as in pair_moments_far:
// each point formed in FP64, as before the cast, inside a rolled 6 x 6
// point loop
template <bool REALK>
__device__ inline void pair_moments32(const double *pa, const double *pb, const float *D, fcx k, int needH1,
                                      Moments32 &m)
{
#pragma clang fp contract(fast)
  m.S1 = fmk(0, 0); m.Sxy = fmk(0, 0); m.s = fmk(0, 0);
  for (int c = 0; c < 3; c++) { m.Sx[c] = fmk(0, 0); m.Sy[c] = fmk(0, 0); m.u[c] = fmk(0, 0); m.v[c] = fmk(0, 0); m.w[c] = fmk(0, 0); }
  const fcx ik = fmk(-k.i, k.r);
#pragma unroll 1
  for (int i = 0; i < 6; i++) {
    const float wi = (float)c_tcr4[3 * i + 2];
    double xd[3];
    tcr4_point(pa, i, xd);
    const float x[3] = {(float)xd[0], (float)xd[1], (float)xd[2]};
#pragma unroll 1
    for (int j = 0; j < 6; j++) {
      const float wj = (float)c_tcr4[3 * j + 2];
      double yd[3];
      tcr4_point(pb, j, yd);
      const float y[3] = {(float)yd[0], (float)yd[1], (float)yd[2]};
      const float R[3] = {D[0] + x[0] - y[0], D[1] + x[1] - y[1], D[2] + x[2] - y[2]};
      const float r = sqrtf(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]), ir = 1.0f / r;
      float sn, cs;
      sincosf(k.r * r, &sn, &cs);
      // (REALK: a real wavenumber, as in a lossless region at real
      // frequency, needs no exponential: expf(-0) is 1 exactly)
      const float e = (REALK ? 1.0f : expf(-k.i * r)) * (float)SCUFFX_INV4PI * ir * (wi * wj);
      const fcx Phi = fmk(e * cs, e * sn);
      m.S1 += Phi;
      for (int c = 0; c < 3; c++) { m.Sx[c] += x[c] * Phi; m.Sy[c] += y[c] * Phi; }
      m.Sxy += (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) * Phi;
      if (needH1) {
        const fcx Psi = (Phi * (ik - fmk(ir, 0))) * fmk(ir, 0);
        const float xy[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
        const float Rx[3] = {R[1] * x[2] - R[2] * x[1], R[2] * x[0] - R[0] * x[2], R[0] * x[1] - R[1] * x[0]};
        const float yR[3] = {y[1] * R[2] - y[2] * R[1], y[2] * R[0] - y[0] * R[2], y[0] * R[1] - y[1] * R[0]};
        m.s += (xy[0] * R[0] + xy[1] * R[1] + xy[2] * R[2]) * Psi;
        for (int c = 0; c < 3; c++) { m.u[c] += Rx[c] * Psi; m.v[c] += yR[c] * Psi; m.w[c] += R[c] * Psi; }
      }
    }
  }
}

__global__ void __launch_bounds__(256)
k_far32(const double *Va, const int *PVIa, const double *PCa, int rowLimit,
        const double *Vb, const int *PVIb, const double *PCb, const int *slotPanels, int same,
        const int *nearRowPtr, const int *nearCol, int nreg, const RegionK *regs,
        int needH1, double *HB, long rowsCap)
{
  const int slot = blockIdx.y;
  const int pa = blockIdx.x * blockDim.x + threadIdx.x;
  if (pa >= rowLimit) return;
  const int pb = slotPanels[slot];
  if (same && pb == pa) return;
  if (near_find(nearRowPtr, nearCol, pa, pb) >= 0) {
    const int stride = needH1 ? 4 : 2;
    HB[((long)slot * rowsCap + pa) * (long)(nreg * 9 * stride)] = __longlong_as_double((long long)SCUFFX_NEAR_MARK);
    return;
  }
  const double *A0 = Va + 3 * PVIa[3 * pa], *A1 = Va + 3 * PVIa[3 * pa + 1], *A2 = Va + 3 * PVIa[3 * pa + 2];
  const double *B0 = Vb + 3 * PVIb[3 * pb], *B1 = Vb + 3 * PVIb[3 * pb + 1], *B2 = Vb + 3 * PVIb[3 * pb + 2];
  double PA[9], PB[9];
  float D[3];
  for (int c = 0; c < 3; c++) {
    const double ca = PCa[4 * pa + c], cb = PCb[4 * pb + c];
    PA[c] = A0[c] - ca; PA[3 + c] = A1[c] - A0[c]; PA[6 + c] = A2[c] - A0[c];
    PB[c] = B0[c] - cb; PB[3 + c] = B1[c] - B0[c]; PB[6 + c] = B2[c] - B0[c];
    D[c] = (float)(ca - cb);
  }
  const int stride = needH1 ? 4 : 2;
  double *out = HB + ((long)slot * rowsCap + pa) * (long)(nreg * 9 * stride);
  for (int r = 0; r < nreg; r++) {
    const cplx k = regs[r].k;
    Moments32 m32;
    const fcx k32 = fmk((float)k.r, (float)k.i);
    if (k32.i == 0) pair_moments32<true>(PA, PB, D, k32, needH1, m32);
    else pair_moments32<false>(PA, PB, D, k32, needH1, m32);
    // the 9 combinations from the single-precision moments
    Moments m;
    m.S1 = mk(m32.S1.r, m32.S1.i); m.Sxy = mk(m32.Sxy.r, m32.Sxy.i); m.s = mk(m32.s.r, m32.s.i);
    for (int c = 0; c < 3; c++) {
      m.Sx[c] = mk(m32.Sx[c].r, m32.Sx[c].i); m.Sy[c] = mk(m32.Sy[c].r, m32.Sy[c].i);
      m.u[c] = mk(m32.u[c].r, m32.u[c].i); m.v[c] = mk(m32.v[c].r, m32.v[c].i); m.w[c] = mk(m32.w[c].r, m32.w[c].i);
    }
    double Qa[9], Qb[9];
    for (int c = 0; c < 3; c++) {
      const double ca = PCa[4 * pa + c], cb = PCb[4 * pb + c];
      Qa[c] = A0[c] - ca; Qa[3 + c] = A1[c] - ca; Qa[6 + c] = A2[c] - ca;
      Qb[c] = B0[c] - cb; Qb[3 + c] = B1[c] - cb; Qb[6 + c] = B2[c] - cb;
    }
    cplx ik = IMUL(k);
    cplx c4 = 4.0 / (ik * ik);
    for (int q = 0; q < 9; q++) {
      cplx h0, h1;
      combo_ab(m, Qa + 3 * (q / 3), Qb + 3 * (q % 3), c4, needH1, &h0, &h1);
      double *o = out + (r * 9 + q) * stride;
      o[0] = h0.r; o[1] = h0.i;
      if (needH1) { o[2] = h1.r; o[3] = h1.i; }
    }
  }
}

hipError_t launch_far(hipStream_t stream, const double *Va, const int *PVIa, const double *PCa, int rowLimit,
                      const double *Vb, const int *PVIb, const double *PCb, const int *slotPanels, int nslots,
                      int same, const int *nearRowPtr, const int *nearCol, int nreg, const RegionK *regs,
                      int needH1, double *HB, long rowsCap, int fp32)
{
  if (rowLimit <= 0 || nslots <= 0 || nreg <= 0) return hipSuccess;
  dim3 grid((rowLimit + 255) / 256, nslots);
  if (fp32)
    hipLaunchKernelGGL(k_far32, grid, dim3(256), 0, stream, Va, PVIa, PCa, rowLimit, Vb, PVIb, PCb, slotPanels, same,
                       nearRowPtr, nearCol, nreg, regs, needH1, HB, rowsCap);
  else
    hipLaunchKernelGGL(k_far, grid, dim3(256), 0, stream, Va, PVIa, PCa, rowLimit, Vb, PVIb, PCb, slotPanels, same,
                       nearRowPtr, nearCol, nreg, regs, needH1, HB, rowsCap);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// K3: near pairs -- desingularized fixed-rule cubature plus the analytic
// singular terms rebuilt from the cached frequency-independent integrals
// (PanelPanelInteractions.cc:504-559, QDFIPPI.cc:67-260).
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(128)
k_near(const NearPair *nps, int nnear, const double *records, int nreg, const RegionK *regs,
       int needH1, double *HN)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnear) return;
  const NearPair &P = nps[t];
  const double *Q = records + (long long)P.rec * SCUFFX_QIF_LEN;
  // reordered vertices: x = Va0 + u(Va1-Va0) + v(Va2-Va0); local origin Va0
  double xi[18], eta[18], Qal[9], Qbl[9], D[3];
  for (int c = 0; c < 3; c++) {
    double aA = P.Va[3 + c] - P.Va[c], aB = P.Va[6 + c] - P.Va[c];
    double bA = P.Vb[3 + c] - P.Vb[c], bB = P.Vb[6 + c] - P.Vb[c];
    for (int i = 0; i < 6; i++) {
      xi[3 * i + c] = c_tcr4[3 * i] * aA + c_tcr4[3 * i + 1] * aB;
      eta[3 * i + c] = c_tcr4[3 * i] * bA + c_tcr4[3 * i + 1] * bB;
    }
    for (int q = 0; q < 3; q++) { Qal[3 * q + c] = P.Qa[3 * q + c] - P.Va[c]; Qbl[3 * q + c] = P.Qb[3 * q + c] - P.Vb[c]; }
    D[c] = P.Va[c] - P.Vb[c];
  }
  // Q-dependent FIPPI recombination, shared pieces
  const double *OVa[3] = {P.OVa0, P.OVa1, P.OVa2}, *OVb[3] = {P.OVb0, P.OVb1, P.OVb2};
  double A[3], B[3], AP[3], BP[3];
  sub3(OVa[1], OVa[0], A); sub3(OVa[2], OVa[1], B);
  sub3(OVb[1], OVb[0], AP); sub3(OVb[2], OVb[1], BP);
  const double *xMxp = Q, *xXxp = Q + 3, *RM1 = Q + 6, *R1 = Q + 15, *R2 = Q + 24, *x0 = Q + 33;
  const double r0m[9] = {1.0 / 4.0, 1.0 / 6.0, 1.0 / 12.0, 1.0 / 6.0, 1.0 / 9.0, 1.0 / 18.0, 1.0 / 12.0, 1.0 / 18.0, 1.0 / 36.0};
  // translation correction (QDFIPPI.cc:143-158): x0 vs centroid of the
  // (possibly swapped) pair as ordered for this call
  double Lc[3];
  const double *Vc = P.flipped ? P.Vb : P.Va;
  for (int c = 0; c < 3; c++) Lc[c] = x0[c] - (Vc[c] + Vc[3 + c] + Vc[6 + c]) / 3.0;
  bool applyL = sqrt(dot3(Lc, Lc)) > 1.0e-8 * sqrt(dot3(A, A));
  double hDot[9][4], hNab[4], hTim[9][4], hTimRM3[9];
  // hDot/hTimes for moments RM1, R0, R1, R2 (QDFIPPI.cc:163-251)
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double *Qa = P.Qa + 3 * a, *Qb = P.Qb + 3 * b;
      const double *OQa = P.flipped ? Qb : Qa, *OQb = P.flipped ? Qa : Qb;
      double V0mQ[3], V0PmQP[3];
      sub3(OVa[0], OQa, V0mQ); sub3(OVb[0], OQb, V0PmQP);
      double VdVP = dot3(V0mQ, V0PmQP), VdAP = dot3(V0mQ, AP), VdBP = dot3(V0mQ, BP);
      double AdVP = dot3(A, V0PmQP), AdAP = dot3(A, AP), AdBP = dot3(A, BP);
      double BdVP = dot3(B, V0PmQP), BdAP = dot3(B, AP), BdBP = dot3(B, BP);
      double QamQb[3], QaxQb[3], V0mV0P[3], S[3];
      sub3(OVa[0], OVb[0], V0mV0P);
      sub3(Qa, Qb, QamQb);
      cross3(Qa, Qb, QaxQb);
      cross3(OVa[0], OVb[0], S); double ht0000 = dot3(QaxQb, V0mV0P) + dot3(QamQb, S);
      cross3(A, OVb[0], S);      double ht1000 = dot3(QaxQb, A) + dot3(QamQb, S);
      cross3(B, OVb[0], S);      double ht0100 = dot3(QaxQb, B) + dot3(QamQb, S);
      cross3(OVa[0], AP, S);     double ht0010 = -dot3(QaxQb, AP) + dot3(QamQb, S);
      cross3(OVa[0], BP, S);     double ht0001 = -dot3(QaxQb, BP) + dot3(QamQb, S);
      cross3(A, AP, S);          double ht1010 = dot3(QamQb, S);
      cross3(A, BP, S);          double ht1001 = dot3(QamQb, S);
      cross3(B, AP, S);          double ht0110 = dot3(QamQb, S);
      cross3(B, BP, S);          double ht0101 = dot3(QamQb, S);
      double rm3 = dot3(QaxQb, xMxp) + dot3(QamQb, xXxp);
      if (applyL) { cross3(Lc, QamQb, S); rm3 -= dot3(S, xMxp); }
      const double *Ms[4] = {RM1, r0m, R1, R2};
      for (int q = 0; q < 4; q++) {
        const double *M = Ms[q];
        hDot[3 * a + b][q] = M[0] * VdVP + M[1] * VdAP + M[2] * VdBP + M[3] * AdVP + M[4] * AdAP
                             + M[5] * AdBP + M[6] * BdVP + M[7] * BdAP + M[8] * BdBP;
        hTim[3 * a + b][q] = M[0] * ht0000 + M[3] * ht1000 + M[6] * ht0100 + M[1] * ht0010
                             + M[2] * ht0001 + M[4] * ht1010 + M[5] * ht1001 + M[7] * ht0110
                             + M[8] * ht0101;
      }
      hTimRM3[3 * a + b] = rm3;
      if (P.flipped) {
        hTimRM3[3 * a + b] *= -1.0;
        for (int q = 0; q < 3; q++) hTim[3 * a + b][q] *= -1.0;
      }
    }
  hNab[0] = 4.0 * RM1[0]; hNab[1] = 1.0; hNab[2] = 4.0 * R1[0]; hNab[3] = 4.0 * R2[0];
  int stride = needH1 ? 4 : 2;
  double *out = HN + (long long)t * (nreg * 9 * stride);
  for (int r = 0; r < nreg; r++) {
    cplx k = regs[r].k;
    Moments m;
    pair_moments<1>(xi, eta, D, k, needH1, m);
    cplx H0[9], H1[9];
    combos_from_moments(m, Qal, Qbl, k, needH1, H0, H1);
    cplx ik = IMUL(k), OOIK2 = 1.0 / (ik * ik), PF[5];
    PF[0] = mk(SCUFFX_INV4PI, 0);
    for (int q = 1; q < 5; q++) PF[q] = ik * PF[q - 1];
    for (int q = 0; q < 9; q++) {
      cplx h0 = H0[q];
      h0 += PF[0] * (mk(hDot[q][0], 0) + OOIK2 * hNab[0])
            + PF[1] * (mk(hDot[q][1], 0) + OOIK2 * hNab[1])
            + PF[2] * 0.5 * (mk(hDot[q][2], 0) + OOIK2 * hNab[2])
            + PF[3] * (1.0 / 6.0) * (mk(hDot[q][3], 0) + OOIK2 * hNab[3]);
      double *o = out + (r * 9 + q) * stride;
      o[0] = h0.r; o[1] = h0.i;
      if (needH1) {
        cplx h1 = H1[q];
        h1 += PF[0] * (-1.0 * hTimRM3[q]) + PF[2] * (0.5 * hTim[q][0])
              + PF[3] * ((1.0 / 3.0) * hTim[q][1]) + PF[4] * ((1.0 / 6.0) * hTim[q][2]);
        o[2] = h1.r; o[3] = h1.i;
      }
    }
  }
}

hipError_t launch_near(hipStream_t stream, const NearPair *np, int nnear, const double *records,
                       int nreg, const RegionK *regs, int needH1, double *HN)
{
  if (nnear <= 0 || nreg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_near, dim3((nnear + 127) / 128), dim3(128), 0, stream, np, nnear, records,
                     nreg, regs, needH1, HN);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// K2: edge pairs of one column band.  GC0 = LaLb(HPP-HPM-HMP+HMM)[0],
// GC1 = LaLb/(ik)(...)[1] (EdgeEdgeInteractions.cc:154-188); EFIE entry
// i w mu s GC0, PMCHWT 2x2 block [[i w mu G, -ik C], [-ik C, -i w eps G]]
// summed over the common regions (SurfaceSurfaceInteractions.cc:162-327).
// grid.y = column edge of the band, threads run down its column (coalesced
// column-major writes); same-surface blocks fill only the upper triangle.
// A panel pair's record comes from the self buffer (a panel with itself),
// the near buffer, or the band's far records (slot of the column panel).
// ---------------------------------------------------------------------
__device__ inline void fetch_h(int pa, int qa, int pb, int qb, int slot, int r, int nreg, int stride, int same,
                               const double *HB, long rowsCap, const double *HN, const double *HS,
                               const int *nearRowPtr, const int *nearCol, const int *nearIdx,
                               cplx *h0, cplx *h1)
{
  const double *rec;
  if (same && pa == pb) {
    rec = HS + ((long long)pa * nreg + r) * 9 * 2 + (3 * qa + qb) * 2;
    *h0 = mk(rec[0], rec[1]);
    *h1 = mk(0, 0);
    return;
  }
  // the band record, unless k_far marked the pair near (its first word)
  const double *band = HB + ((long)slot * rowsCap + pa) * (long)(nreg * 9 * stride);
  if ((unsigned long long)__double_as_longlong(band[0]) == SCUFFX_NEAR_MARK) {
    int ni = near_find(nearRowPtr, nearCol, pa, pb);
    int id = ni >= 0 ? nearIdx[ni] : -1;
    if (id < 0) {   // a near pair outside the upper-triangle walk: never requested
      *h0 = mk(0, 0); *h1 = mk(0, 0);
      return;
    }
    rec = HN + ((long long)id * nreg + r) * 9 * stride + (3 * qa + qb) * stride;
  } else {
    rec = band + (r * 9 + 3 * qa + qb) * stride;
  }
  *h0 = mk(rec[0], rec[1]);
  *h1 = stride == 4 ? mk(rec[2], rec[3]) : mk(0, 0);
}

__global__ void __launch_bounds__(256)
k_stamp(int nEa, const int *EIa, const double *ELa, const int *bandCols, const int *colSlots, const int *EIb,
        const double *ELb, int same, int pecA, int pecB, int nreg, const RegionK *regs, int needH1,
        const double *HB, long rowsCap, const double *HN, const double *HS,
        const int *nearRowPtr, const int *nearCol, const int *nearIdx,
        MatOut M, int rowOff, int colOff)
{
  const int ea = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (ea >= nEa) return;
  const int eb = bandCols[j];
  if (same && ea > eb) return;
  int stride = needH1 ? 4 : 2;
  int pA[2] = {EIa[4 * ea], EIa[4 * ea + 1]}, qA[2] = {EIa[4 * ea + 2], EIa[4 * ea + 3]};
  int pB[2] = {EIb[4 * eb], EIb[4 * eb + 1]}, qB[2] = {EIb[4 * eb + 2], EIb[4 * eb + 3]};
  int sB[2] = {colSlots[2 * j], colSlots[2 * j + 1]};
  double LL = ELa[ea] * ELb[eb];
  cplx m00 = mk(0, 0), m01 = mk(0, 0), m11 = mk(0, 0);
  for (int r = 0; r < nreg; r++) {
    cplx G[4], C[4];
    // the four far records and their near marks loaded together, from
    // valid addresses whatever the pair turns out to be (a load behind the
    // mark test would wait for it, four times per region); self and near
    // pairs, the exceptions, then read their own buffers
    double mw[4], rv[4][4];
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const int s = st >> 1, t = st & 1;
      const bool ok = pA[s] >= 0 && pB[t] >= 0;
      const double *band = HB + ((long)(ok ? sB[t] : 0) * rowsCap + (ok ? pA[s] : 0)) * (long)(nreg * 9 * stride);
      const double *rec = band + (r * 9 + (ok ? 3 * qA[s] + qB[t] : 0)) * stride;
      mw[st] = band[0];
      rv[st][0] = rec[0]; rv[st][1] = rec[1];
      if (stride == 4) { rv[st][2] = rec[2]; rv[st][3] = rec[3]; }   // (records of 2 doubles end at rec[1])
      else { rv[st][2] = 0.0; rv[st][3] = 0.0; }
    }
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const int s = st >> 1, t = st & 1;
      G[st] = mk(0, 0); C[st] = mk(0, 0);
      if (pA[s] < 0 || pB[t] < 0) continue;
      if ((same && pA[s] == pB[t]) || (unsigned long long)__double_as_longlong(mw[st]) == SCUFFX_NEAR_MARK)
        fetch_h(pA[s], qA[s], pB[t], qB[t], sB[t], r, nreg, stride, same, HB, rowsCap, HN, HS,
                nearRowPtr, nearCol, nearIdx, &G[st], &C[st]);
      else {
        G[st] = mk(rv[st][0], rv[st][1]);
        C[st] = stride == 4 ? mk(rv[st][2], rv[st][3]) : mk(0, 0);
      }
    }
    cplx k = regs[r].k;
    cplx GC0 = LL * (G[0] - G[1] - G[2] + G[3]);
    cplx GC1 = (mk(LL, 0) / IMUL(k)) * (C[0] - C[1] - C[2] + C[3]);
    if (k.r == 0.0 && k.i == 0.0) { GC0 = mk(0, 0); GC1 = mk(0, 0); }
    m00 += regs[r].pf1 * GC0;
    m01 += regs[r].pf2 * GC1;
    m11 += regs[r].pf3 * GC0;
  }
  long R = rowOff, Cc = colOff;
  if (pecA && pecB) {
    M.put(R + ea, Cc + eb, m00);
  } else if (pecA) {
    M.put(R + ea, Cc + 2 * eb, m00);
    M.put(R + ea, Cc + 2 * eb + 1, m01);
  } else if (pecB) {
    M.put(R + 2 * ea, Cc + eb, m00);
    M.put(R + 2 * ea + 1, Cc + eb, m01);
  } else {
    long x = R + 2 * ea, y = Cc + 2 * eb;
    M.put(x, y, m00);
    if (!same || ea != eb) M.put(x + 1, y, m01);
    M.put(x, y + 1, m01);
    M.put(x + 1, y + 1, m11);
  }
}

hipError_t launch_stamp(hipStream_t stream, int nEa, const int *EIa, const double *ELa, const int *bandCols,
                        const int *colSlots, int ncols, int rowEdges, const int *EIb, const double *ELb, int same,
                        int pecA, int pecB, int nreg, const RegionK *regs, int needH1,
                        const double *HB, long rowsCap, const double *HN, const double *HS,
                        const int *nearRowPtr, const int *nearCol, const int *nearIdx,
                        MatOut M, int rowOff, int colOff)
{
  if (nEa <= 0 || ncols <= 0 || rowEdges <= 0) return hipSuccess;
  dim3 grid((std::min(nEa, rowEdges) + 255) / 256, ncols);
  hipLaunchKernelGGL(k_stamp, grid, dim3(256), 0, stream, nEa, EIa, ELa, bandCols, colSlots, EIb, ELb, same, pecA,
                     pecB, nreg, regs, needH1, HB, rowsCap, HN, HS, nearRowPtr, nearCol, nearIdx, M, rowOff,
                     colOff);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// lower triangle <- transpose of the upper triangle
// (AssembleBEMMatrix.cc:753-758), 32x32 tiles staged through LDS
// ---------------------------------------------------------------------
template <typename E>
__global__ void __launch_bounds__(256) k_mirror(E *M, long ldm, int n)
{
  __shared__ E tile[32][33];
  int bi = blockIdx.y, bj = blockIdx.x;   // destination tile (rows bi, cols bj), bi >= bj
  if (bj > bi) return;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  // read source tile: rows bj*32.., cols bi*32.. (upper part)
  for (int yy = ty; yy < 32; yy += 8) {
    int r = bj * 32 + tx, c = bi * 32 + yy;
    if (r < n && c < n) tile[yy][tx] = M[r + (long)c * ldm];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int r = bi * 32 + tx, c = bj * 32 + yy;
    if (r < n && c < n && r > c) M[r + (long)c * ldm] = tile[tx][yy];
  }
}

hipError_t launch_mirror(hipStream_t stream, MatOut M, int n)
{
  if (n <= 1 || M.kind == MAT_F64_PACKED) return hipSuccess;   // packed storage holds the upper triangle only
  int nt = (n + 31) / 32;
  if (M.kind == MAT_C128)
    hipLaunchKernelGGL(k_mirror<double2>, dim3(nt, nt), dim3(256), 0, stream, (double2 *)M.p, M.ld, n);
  else
    hipLaunchKernelGGL(k_mirror<double>, dim3(nt, nt), dim3(256), 0, stream, (double *)M.p, M.ld, n);
  return hipGetLastError();
}

// diagonal block of surface size n copied from offset `from` to offset `to`
// of a packed upper-triangular matrix (the Mate copy of packed storage)
__global__ void k_copy_packed_block(double *M, long from, long to, long n)
{
  const long c = blockIdx.y, r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r > c || c >= n) return;
  const long cf = from + c, ct = to + c;
  M[(to + r) + ct * (ct + 1) / 2] = M[(from + r) + cf * (cf + 1) / 2];
}

hipError_t launch_copy_block(hipStream_t stream, MatOut M, long from, long to, long n)
{
  if (n <= 0) return hipSuccess;
  if (M.kind == MAT_F64_PACKED) {
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(k_copy_packed_block, grid, dim3(256), 0, stream, (double *)M.p, from, to, n);
    return hipGetLastError();
  }
  const size_t es = M.kind == MAT_C128 ? sizeof(double2) : sizeof(double);
  char *base = (char *)M.p;
  return hipMemcpy2DAsync(base + es * (to + to * M.ld), es * M.ld, base + es * (from + from * M.ld), es * M.ld,
                          es * n, n, hipMemcpyDeviceToDevice, stream);
}

}  // namespace scuffx
