// kern_grad.hip -- displacement derivatives of a BEM block between two
// distinct surfaces (GradM of AssembleBEMMatrixBlock).
//
// For every RWG pair the four panel-pair integrals are differentiated with
// respect to a rigid displacement of the row surface
// (PanelPanelInteractions.cc:132-199, :176-180):
//   dH0/dx_m = sum w w' R_m (F.F' + 4/(ik)^2) Psi
//   dH1/dx_m = sum w w' [R_m ((F x F').R) Zeta + (F x F')_m Psi]
// with Phi = e^{ikr}/4 pi r, Psi = Phi (ik - 1/r)/r,
// Zeta = Phi ((ik)^2 - 3ik/r + 3/r^2)/r^2, R = x - x'.  Panel pairs with
// relative distance > 4 use the 6-point rule, nearer ones the 78-point
// rule without desingularization, as the reference does when gradients
// are requested (:504-510).  The edge-pair combination and the region
// prefactors follow EdgeEdgeInteractions.cc:183-198 and
// SurfaceSurfaceInteractions.cc:212-327.
#include <hip/hip_runtime.h>

#include "bem.h"
#include "tcr20.inc"

namespace scuffx {

__constant__ double c_gtcr4[18] = SCUFFX_TCR4_INIT;
__constant__ double c_gtcr20[234];

#define SCUFFX_INV4PI_G 0.079577471545947667884

// gradient of one panel-pair integral: G[2m] (H0 part), G[2m+1] (H1 part);
// angular derivatives about the GS.n axes: T[2t], T[2t+1]
// (PanelPanelInteractions.cc:182-199: dX = Gamma X, dF = Gamma F,
//  dH0 += hPlus (R.dX) Psi + (dF.F') Phi,
//  dH1 += hTimes (R.dX) Zeta + ((dF x F').R + (F x F').dX) Psi)
__device__ void ppi_grad(const double *Va0, const double *Va1, const double *Va2, const double *Qa,
                         const double *Vb0, const double *Vb1, const double *Vb2, const double *Qb,
                         cplx k, int high, const GammaSet &GS, cplx G[6], cplx T[6])
{
  double A[3], B[3], AP[3], BP[3];
  sub3(Va1, Va0, A); sub3(Va2, Va0, B);
  sub3(Vb1, Vb0, AP); sub3(Vb2, Vb0, BP);
  const double *rule = high ? c_gtcr20 : c_gtcr4;
  const int np = high ? 78 : 6;
  cplx ik = IMUL(k), ik2 = ik * ik, c4 = 4.0 / ik2;
  for (int m = 0; m < 6; m++) { G[m] = mk(0, 0); T[m] = mk(0, 0); }
  const int ntq = GS.n;
  for (int i = 0; i < np; i++) {
    const double u = rule[3 * i], v = rule[3 * i + 1], w = rule[3 * i + 2];
    double X[3], F[3];
    for (int c = 0; c < 3; c++) { X[c] = Va0[c] + u * A[c] + v * B[c]; F[c] = X[c] - Qa[c]; }
    // rotated position and source vector of the outer point, per axis
    double dX[3][3], dF[3][3];
    for (int t = 0; t < ntq; t++)
      for (int m = 0; m < 3; m++) {
        dX[t][m] = 0; dF[t][m] = 0;
        for (int q = 0; q < 3; q++) {
          dX[t][m] += GS.g[9 * t + m + 3 * q] * X[q];
          dF[t][m] += GS.g[9 * t + m + 3 * q] * F[q];
        }
      }
    cplx GI[6], TI[6];
    for (int m = 0; m < 6; m++) { GI[m] = mk(0, 0); TI[m] = mk(0, 0); }
    for (int j = 0; j < np; j++) {
      const double up = rule[3 * j], vp = rule[3 * j + 1], wp = rule[3 * j + 2];
      double FP[3], R[3];
      for (int c = 0; c < 3; c++) {
        double XP = Vb0[c] + up * AP[c] + vp * BP[c];
        FP[c] = XP - Qb[c];
        R[c] = X[c] - XP;
      }
      double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2], r = sqrt(r2), ir = 1.0 / r;
      double e = exp(-k.i * r), sn, cs;
      sincos(k.r * r, &sn, &cs);
      cplx Phi = mk(e * cs, e * sn) * (SCUFFX_INV4PI_G * ir * wp);
      cplx Psi = Phi * (ik - mk(ir, 0)) * ir;
      cplx Zeta = Phi * (ik2 - 3.0 * ir * ik + mk(3.0 * ir * ir, 0)) * (ir * ir);
      double FxFP[3];
      cross3(F, FP, FxFP);
      cplx hPlus = mk(F[0] * FP[0] + F[1] * FP[1] + F[2] * FP[2], 0) + c4;
      double hTimes = FxFP[0] * R[0] + FxFP[1] * R[1] + FxFP[2] * R[2];
      cplx hPsi = hPlus * Psi, tZeta = hTimes * Zeta;
      for (int m = 0; m < 3; m++) {
        GI[2 * m] += R[m] * hPsi;
        GI[2 * m + 1] += R[m] * tZeta + FxFP[m] * Psi;
      }
      for (int t = 0; t < ntq; t++) {
        double Puv = R[0] * dX[t][0] + R[1] * dX[t][1] + R[2] * dX[t][2];
        double dFFP = dF[t][0] * FP[0] + dF[t][1] * FP[1] + dF[t][2] * FP[2];
        double dFxFP[3];
        cross3(dF[t], FP, dFxFP);
        double c1 = dFxFP[0] * R[0] + dFxFP[1] * R[1] + dFxFP[2] * R[2]
                    + FxFP[0] * dX[t][0] + FxFP[1] * dX[t][1] + FxFP[2] * dX[t][2];
        TI[2 * t] += Puv * hPsi + dFFP * Phi;
        TI[2 * t + 1] += Puv * tZeta + c1 * Psi;
      }
    }
    for (int m = 0; m < 6; m++) { G[m] += w * GI[m]; T[m] += w * TI[m]; }
  }
}

// one thread per (ea, eb) of the block; GB holds three column-major
// NBF_a x NBF_b matrices (leading dimension ldg, gstride elements apart),
// overwritten
__global__ void __launch_bounds__(128)
k_grad_stamp(int nEa, const int *EIa, const double *ELa, const double *Va, const int *PVIa, const double *PCa,
             int nEb, const int *EIb, const double *ELb, const double *Vb, const int *PVIb, const double *PCb,
             int pecA, int pecB, int nreg, const RegionK *regs, double2 *GB, long ldg, long gstride,
             GammaSet GS, double2 *TB)
{
  int ea = blockIdx.x * blockDim.x + threadIdx.x;
  int eb = blockIdx.y;
  if (ea >= nEa || eb >= nEb) return;
  int pA[2] = {EIa[4 * ea], EIa[4 * ea + 1]}, qA[2] = {EIa[4 * ea + 2], EIa[4 * ea + 3]};
  int pB[2] = {EIb[4 * eb], EIb[4 * eb + 1]}, qB[2] = {EIb[4 * eb + 2], EIb[4 * eb + 3]};
  double LL = ELa[ea] * ELb[eb];
  const int nd = 3 + GS.n;           // displacement components, then axes
  cplx m00[6], m01[6], m11[6];
  for (int m = 0; m < 6; m++) { m00[m] = mk(0, 0); m01[m] = mk(0, 0); m11[m] = mk(0, 0); }
  for (int r = 0; r < nreg; r++) {
    cplx k = regs[r].k;
    cplx S0[6], S1[6];
    for (int m = 0; m < 6; m++) { S0[m] = mk(0, 0); S1[m] = mk(0, 0); }
    for (int s = 0; s < 2; s++)
      for (int t = 0; t < 2; t++) {
        if (pA[s] < 0 || pB[t] < 0) continue;
        const int pa = pA[s], pb = pB[t];
        const double *a0 = Va + 3 * PVIa[3 * pa], *a1 = Va + 3 * PVIa[3 * pa + 1], *a2 = Va + 3 * PVIa[3 * pa + 2];
        const double *b0 = Vb + 3 * PVIb[3 * pb], *b1 = Vb + 3 * PVIb[3 * pb + 1], *b2 = Vb + 3 * PVIb[3 * pb + 2];
        const double *Qa = Va + 3 * PVIa[3 * pa + qA[s]], *Qb = Vb + 3 * PVIb[3 * pb + qB[t]];
        // relative centroid distance (AssessPanelPair.cc:161-166)
        double dc[3] = {PCa[4 * pa] - PCb[4 * pb], PCa[4 * pa + 1] - PCb[4 * pb + 1], PCa[4 * pa + 2] - PCb[4 * pb + 2]};
        double rRel = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) / fmax(PCa[4 * pa + 3], PCb[4 * pb + 3]);
        cplx G[6], TT[6];
        ppi_grad(a0, a1, a2, Qa, b0, b1, b2, Qb, k, rRel > 4.0 ? 0 : 1, GS, G, TT);
        double sg = ((s + t) & 1) ? -1.0 : 1.0;
        for (int m = 0; m < 3; m++) { S0[m] += sg * G[2 * m]; S1[m] += sg * G[2 * m + 1]; }
        for (int q = 0; q < GS.n; q++) { S0[3 + q] += sg * TT[2 * q]; S1[3 + q] += sg * TT[2 * q + 1]; }
      }
    cplx cpre = mk(LL, 0) / IMUL(k);
    for (int m = 0; m < nd; m++) {
      cplx g0 = LL * S0[m], g1 = cpre * S1[m];
      m00[m] += regs[r].pf1 * g0;
      m01[m] += regs[r].pf2 * g1;
      m11[m] += regs[r].pf3 * g0;
    }
  }
  for (int m = 0; m < nd; m++) {
    double2 *T = (m < 3) ? GB + (long)m * gstride : TB + (long)(m - 3) * gstride;
    if (pecA && pecB) {
      T[ea + (long)eb * ldg] = make_double2(m00[m].r, m00[m].i);
    } else if (pecA) {
      T[ea + (long)(2 * eb) * ldg] = make_double2(m00[m].r, m00[m].i);
      T[ea + (long)(2 * eb + 1) * ldg] = make_double2(m01[m].r, m01[m].i);
    } else if (pecB) {
      T[2 * ea + (long)eb * ldg] = make_double2(m00[m].r, m00[m].i);
      T[2 * ea + 1 + (long)eb * ldg] = make_double2(m01[m].r, m01[m].i);
    } else {
      long x = 2 * ea, y = 2 * eb;
      T[x + y * ldg] = make_double2(m00[m].r, m00[m].i);
      T[(x + 1) + y * ldg] = make_double2(m01[m].r, m01[m].i);
      T[x + (y + 1) * ldg] = make_double2(m01[m].r, m01[m].i);
      T[(x + 1) + (y + 1) * ldg] = make_double2(m11[m].r, m11[m].i);
    }
  }
}

static bool g_gtcr20_ready[64] = {false};

hipError_t launch_grad(hipStream_t stream, int nEa, const int *EIa, const double *ELa, const double *Va,
                       const int *PVIa, const double *PCa, int nEb, const int *EIb, const double *ELb,
                       const double *Vb, const int *PVIb, const double *PCb, int pecA, int pecB, int nreg,
                       const RegionK *regs, double2 *GB, long ldg, long gstride, const GammaSet &GS,
                       double2 *TB)
{
  int dev = 0;
  hipGetDevice(&dev);
  if (!g_gtcr20_ready[dev & 63]) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_gtcr20), scuffx_tcr20, sizeof(scuffx_tcr20));
    if (e != hipSuccess) return e;
    g_gtcr20_ready[dev & 63] = true;
  }
  if (nEa <= 0 || nEb <= 0) return hipSuccess;
  dim3 grid((nEa + 127) / 128, nEb);
  hipLaunchKernelGGL(k_grad_stamp, grid, dim3(128), 0, stream, nEa, EIa, ELa, Va, PVIa, PCa, nEb, EIb, ELb, Vb,
                     PVIb, PCb, pecA, pecB, nreg, regs, GB, ldg, gstride, GS, TB);
  return hipGetLastError();
}

}  // namespace scuffx
