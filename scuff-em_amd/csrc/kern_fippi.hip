// kern_fippi.hip -- frequency-independent panel-panel integrals (FIPPIs)
// by h-adaptive cubature, one wavefront per unique panel-pair job.
//
// Restates the serial h-adaptive scheme of libs/libSGJC/hcubature.c
// (:455-615 Genz-Malik 7/5 rule for 2-D/3-D, :619-758 Gauss-Kronrod 15
// for 1-D, :772-846 max-heap on the region error, :887-931 driver with
// ERROR_INDIVIDUAL convergence) applied to
//   ncv = 0: the 3-D integrand of QIFIPPICubature.cc:187-325
//            (ABSTOL 1e-10, RELTOL 1e-6, MAXEVALS 1e6)
//   ncv > 0: the 1-D (common edge) / 2-D (common vertex) Taylor-Duffy
//            integrand of QIFIPPITaylorDuffyV2P0.cc:91-221, :329-900
//            (ABSTOL 0, RELTOL 1e-5, MAXFEVALS 1e4)
// Per iteration the lanes evaluate the 2 x (15|17|33) rule points of
// the two halves of the split region in parallel into LDS; lanes 0..32
// then form each function's rule sums, lane 0 keeps the heap (in HBM,
// one slab per wave slot).  Heap order, running totals and the final
// re-summation follow the reference exactly.
#include <hip/hip_runtime.h>

#include "bem.h"

namespace scuffx {

#define NF 33
#define REC 76                 // center[3] half[3] vol split val[33] err[33] pad

struct FipW {
  // ncv = 0
  double V0[3], A[3], B[3], V0P[3], AP[3], BP[3];
  // ncv > 0
  double tA[3], tB[3], tAP[3], tBP[3];
  double V0xAP[3], V0xBP[3], AxV0[3], BxV0[3], AxAP[3], AxBP[3], BxAP[3], BxBP[3];
  double A2, B2, AP2, BP2, L2, AdB, AdAP, AdBP, AdL, BdAP, BdBP, APdBP, BPdL;
  int ncv, ce;
};

// ---------------------------- ncv = 0 --------------------------------
__device__ void cfd3(const FipW &W, const double *x, double *fval)
{
  double u = x[0], v = u * x[1], up = x[2];
  double X[3], Y[3], Z[3], BPdY = 0, Y2 = 0, a2 = 0;
  for (int m = 0; m < 3; m++) {
    X[m] = W.V0[m] + u * W.A[m] + v * W.B[m];
    Z[m] = W.V0P[m] + up * W.AP[m];
    Y[m] = X[m] - Z[m];
    a2 += W.BP[m] * W.BP[m];
    BPdY += W.BP[m] * Y[m];
    Y2 += Y[m] * Y[m];
  }
  double a = sqrt(a2), a3 = a * a2;
  double vp0 = -BPdY / a2, vp02 = vp0 * vp0, vp03 = vp02 * vp0, vp04 = vp03 * vp0;
  double b2 = Y2 / a2 - vp02;
  double OneRM3, vpRM3, OneRM1, vpRM1, OneR1, vpR1, OneR2, vpR2;
  if (b2 > 1.0e-10) {
    double S1 = sqrt(b2 + vp02);
    double S2 = sqrt(b2 + (vp0 + up) * (vp0 + up));
    double LogFac = log((S2 + (up + vp0)) / (S1 + vp0));
    double Sum = vp0 + up, Sum3 = Sum * Sum * Sum, Sum4 = Sum3 * Sum;
    OneRM3 = u * ((up + vp0) / S2 - vp0 / S1) / (a3 * b2);
    vpRM3 = -vp0 * OneRM3 + u * (1.0 / S1 - 1.0 / S2) / a3;
    OneRM1 = u * LogFac / a;
    vpRM1 = -vp0 * OneRM1 + u * (S2 - S1) / a;
    OneR1 = u * a * 0.5 * (b2 * LogFac + (up + vp0) * S2 - vp0 * S1);
    vpR1 = -vp0 * OneR1 + u * a * (S2 * S2 * S2 - S1 * S1 * S1) / 3.0;
    OneR2 = u * a2 * ((Sum3 - vp03) / 3.0 + up * b2);
    vpR2 = -vp0 * OneR2 + u * a2 * ((Sum4 - vp04) / 4.0 + up * (vp0 + 0.5 * up) * b2);
  } else {
    double s = vp0 + up, s2 = s * s, s3 = s2 * s, s4 = s3 * s;
    OneRM3 = u * fabs(1.0 / vp02 - 1.0 / s2) / (2.0 * a3);
    vpRM3 = -vp0 * OneRM3 + u * fabs(1.0 / vp0 - 1.0 / s) / a3;
    OneRM1 = u * fabs(log(s / vp0)) / a;
    vpRM1 = -vp0 * OneRM1 + u * up / a;
    OneR1 = u * a * fabs(s2 - vp02) / 2.0;
    OneR2 = u * a2 * fabs(s3 - vp03) / 3.0;
    vpR1 = -vp0 * OneR1 + OneR2 / a;
    vpR2 = -vp0 * OneR2 + u * a2 * fabs(s4 - vp04) / 4.0;
  }
  double CPCT[3], CPLT[3];
  cross3(X, Z, CPCT);
  cross3(X, W.BP, CPLT);
  int nf = 0;
  for (int m = 0; m < 3; m++) fval[nf++] = Y[m] * OneRM3 - W.BP[m] * vpRM3;
  for (int m = 0; m < 3; m++) fval[nf++] = CPCT[m] * OneRM3 + CPLT[m] * vpRM3;
  const double I1[3] = {OneRM1, OneR1, OneR2}, Iv[3] = {vpRM1, vpR1, vpR2};
  for (int q = 0; q < 3; q++) {
    fval[nf++] = I1[q];
    fval[nf++] = up * I1[q];
    fval[nf++] = Iv[q];
    fval[nf++] = u * I1[q];
    fval[nf++] = u * up * I1[q];
    fval[nf++] = u * Iv[q];
    fval[nf++] = v * I1[q];
    fval[nf++] = v * up * I1[q];
    fval[nf++] = v * Iv[q];
  }
}

// ---------------------------- ncv > 0 --------------------------------
__device__ inline void pqr_abg2(double P, double Q, double R, double *A, double *B, double *G2)
{
  *A = sqrt(P);
  *B = Q / P;
  double RoP = R / P;
  *G2 = RoP - (*B) * (*B);
  if (fabs(*G2) < 1.0e-6 * RoP) { *B = 0.0; *G2 = RoP; }
}

__device__ void tdf_abg(const FipW &W, const double *y, double *A, double *B, double *G2)
{
  if (W.ce) {
    double y1 = y[0], y12 = y1 * y1;
    double A2 = W.A2, BP2 = W.BP2, L2 = W.L2, AdBP = W.AdBP, AdL = W.AdL, BPdL = W.BPdL;
    pqr_abg2((BP2 - 2.0 * BPdL + L2) * y12, y1 * (L2 + BPdL * (y1 - 1.0) - (AdL + L2 - AdBP) * y1),
             L2 - 2 * AdL * y1 - 2 * L2 * y1 + A2 * y12 + 2 * AdL * y12 + L2 * y12, A + 0, B + 0, G2 + 0);
    pqr_abg2(BP2 * y12, (BPdL + AdBP * y1 - BPdL * y1) * y1,
             L2 + 2 * AdL * y1 - 2 * L2 * y1 + (A2 - 2 * AdL + L2) * y12, A + 1, B + 1, G2 + 1);
    pqr_abg2((A2 + 2 * AdBP + BP2) * y12, y1 * (AdL * (-1.0 + y1) + BPdL * (-1.0 + y1) - (AdBP + BP2) * y1),
             L2 + 2 * BPdL * y1 - 2 * L2 * y1 + BP2 * y12 - 2 * BPdL * y12 + L2 * y12, A + 2, B + 2, G2 + 2);
    pqr_abg2((A2 + 2 * AdBP - 2 * AdL + BP2 - 2 * BPdL + L2) * y12,
             y1 * (AdL - L2 - (AdBP + BP2) * y1 + BPdL * (1 + y1)), L2 - 2 * BPdL * y1 + BP2 * y12,
             A + 3, B + 3, G2 + 3);
    pqr_abg2(A2 * y12, (AdBP * y1 - AdL) * y1, L2 - 2 * BPdL * y1 + BP2 * y12, A + 4, B + 4, G2 + 4);
    pqr_abg2(A2 * y12, y1 * (AdL + AdBP * y1 - AdL * y1),
             L2 + 2 * BPdL * y1 - 2 * L2 * y1 + BP2 * y12 - 2 * BPdL * y12 + L2 * y12, A + 5, B + 5, G2 + 5);
  } else {
    double y1 = y[0], y2 = y[1], y22 = y2 * y2;
    double A2 = W.A2, B2 = W.B2, AP2 = W.AP2, BP2 = W.BP2, AdB = W.AdB, AdAP = W.AdAP, AdBP = W.AdBP,
           BdAP = W.BdAP, BdBP = W.BdBP, APdBP = W.APdBP;
    pqr_abg2(BP2 * y22, -AdBP * y2 - BdBP * y1 * y2 + APdBP * y22,
             A2 + B2 * y1 * y1 + AP2 * y2 * y2 + 2.0 * AdB * y1 - 2.0 * AdAP * y2 - 2.0 * BdAP * y1 * y2,
             A + 0, B + 0, G2 + 0);
    pqr_abg2(B2 * y22, AdB * y22 - BdAP * y2 - BdBP * y1 * y2,
             A2 * y2 * y2 + AP2 + BP2 * y1 * y1 - 2 * AdAP * y2 - 2 * AdBP * y2 * y1 + 2 * APdBP * y1,
             A + 1, B + 1, G2 + 1);
  }
}

__device__ inline void tdf_qfp(double P, double Q2, double *Iq, double *Iy)
{
  double P2 = P * P, PP1 = P + 1.0;
  if (fabs(Q2) < 1.0e-8) {
    double SignP = P > 0.0 ? 1.0 : -1.0;
    Iq[0] = SignP * (P + 0.5) / (P2 * PP1 * PP1);
    Iy[0] = -P * Iq[0] + SignP / (P * PP1);
    Iq[1] = SignP * log(1.0 + 1.0 / P);
    Iy[1] = -P * Iq[1] + SignP * 1.0;
    Iq[2] = SignP * (P + 0.5);
    Iy[2] = -P * Iq[2] + SignP * (P * PP1 + 1.0 / 3.0);
    Iq[3] = P * PP1 + 1.0 / 3.0;
    Iy[3] = -P * Iq[3] + P * P2 + (3.0 / 2.0) * P2 + P + 0.25;
  } else {
    double S2 = P * P + Q2, S = sqrt(S2), S3 = S * S2, S4 = S2 * S2;
    double T2 = PP1 * PP1 + Q2, T = sqrt(T2), T3 = T * T2, T4 = T2 * T2;
    double LogFac = log((PP1 + T) / (P + S));
    Iq[0] = (PP1 / T - P / S) / Q2;
    Iy[0] = -P * Iq[0] + 1.0 / S - 1.0 / T;
    Iq[1] = LogFac;
    Iy[1] = -P * Iq[1] + (T - S);
    Iq[2] = (T + P * (T - S) + Q2 * LogFac) / 2.0;
    Iy[2] = -P * Iq[2] + (T3 - S3) / 3.0;
    Iq[3] = (T2 + S2) / 2.0 - 1.0 / 6.0;
    Iy[3] = -P * Iq[3] + (T4 - S4) / 4.0;
  }
}

// polynomial tables of the Duffy-transformed numerators; see the oracle
// restatement for the slot layout (PCross[slot][region][n][j])
__device__ void tdf_tables_ce(const FipW &W, double y1, double PC[6][6][4][2], double PX[9][6][4][2])
{
  const double *A = W.tA, *B = W.tB, *AP = W.tAP, *BP = W.tBP;
  const double *V0xAP = W.V0xAP, *V0xBP = W.V0xBP, *AxV0 = W.AxV0, *BxV0 = W.BxV0;
  const double *AxAP = W.AxAP, *AxBP = W.AxBP, *BxAP = W.BxAP, *BxBP = W.BxBP;
  for (int m = 0; m < 3; m++) {
    PC[m][0][1][0] = (BP[m]-B[m]+AP[m])*y1-BP[m]+B[m];
    PC[m][0][1][1] = B[m]*y1;
    PC[m][0][2][0] = -((2.0*BP[m]-2.0*B[m]+2.0*AP[m])*y1-2.0*BP[m]+2.0*B[m]-AP[m]+A[m])/2.0;
    PC[m][0][2][1] = -B[m]*y1;
    PC[m][0][3][0] = 0.0; PC[m][0][3][1] = 0.0;
    PC[m][1][1][0] = (BP[m]-B[m]-A[m])*y1-BP[m]+B[m];
    PC[m][1][1][1] = -BP[m]*y1;
    PC[m][1][2][0] = ((-2.0*BP[m]+2.0*B[m]+2.0*A[m])*y1+2.0*BP[m]-2.0*B[m]+AP[m]-A[m])/2.0;
    PC[m][1][2][1] = BP[m]*y1;
    PC[m][1][3][0] = 0.0; PC[m][1][3][1] = 0.0;
    PC[m][2][1][0] = -B[m]*y1-BP[m]+B[m];
    PC[m][2][1][1] = (BP[m]+AP[m])*y1;
    PC[m][2][2][0] = -(-2.0*B[m]*y1-2.0*BP[m]+2.0*B[m]-AP[m]+A[m])/2.0;
    PC[m][2][2][1] = -(BP[m]+AP[m])*y1;
    PC[m][2][3][0] = 0.0; PC[m][2][3][1] = 0.0;
    PC[m][3][1][0] = BP[m]*y1-BP[m]+B[m];
    PC[m][3][1][1] = (-B[m]-A[m])*y1;
    PC[m][3][2][0] = (-2.0*BP[m]*y1+2.0*BP[m]-2.0*B[m]+AP[m]-A[m])/2.0;
    PC[m][3][2][1] = (B[m]+A[m])*y1;
    PC[m][3][3][0] = 0.0; PC[m][3][3][1] = 0.0;
    PC[m][4][1][0] = BP[m]*y1-BP[m]+B[m];
    PC[m][4][1][1] = AP[m]*y1;
    PC[m][4][2][0] = -(2.0*BP[m]*y1-2.0*BP[m]+2.0*B[m]-AP[m]+A[m])/2.0;
    PC[m][4][2][1] = -AP[m]*y1;
    PC[m][4][3][0] = 0.0; PC[m][4][3][1] = 0.0;
    PC[m][5][1][0] = -B[m]*y1-BP[m]+B[m];
    PC[m][5][1][1] = -A[m]*y1;
    PC[m][5][2][0] = (2.0*B[m]*y1+2.0*BP[m]-2.0*B[m]+AP[m]-A[m])/2.0;
    PC[m][5][2][1] = A[m]*y1;
    PC[m][5][3][0] = 0.0; PC[m][5][3][1] = 0.0;

    PC[m+3][0][1][0] = ((-2.0*V0xBP[m]-2.0*V0xAP[m]-2.0*BxV0[m]-BxAP[m]-AxBP[m]-AxAP[m])*y1
                        +2.0*V0xBP[m]+2.0*BxV0[m]+BxAP[m]+AxBP[m])/2.0;
    PC[m+3][0][1][1] = (2.0*BxV0[m]+BxAP[m])*y1/2.0;
    PC[m+3][0][2][0] = -((-2.0*BxBP[m]-2.0*BxAP[m])*y1*y1
                         +(-2.0*V0xBP[m]-2.0*V0xAP[m]-2.0*BxV0[m]+4*BxBP[m]+2.0*BxAP[m])*y1
                         +2.0*V0xBP[m]+V0xAP[m]+2.0*BxV0[m]-2.0*BxBP[m]+AxV0[m])/2.0;
    PC[m+3][0][2][1] = -(((2.0*BxBP[m]+2.0*BxAP[m])*y1*y1+(2.0*BxV0[m]-2.0*BxBP[m])*y1))/2.0;
    PC[m+3][0][3][0] = ((-6.0*BxBP[m]-6.0*BxAP[m])*y1*y1
                        +(12.0*BxBP[m]+9*BxAP[m]+3.0*AxBP[m]+3.0*AxAP[m])*y1-6.0*BxBP[m]
                        -3.0*BxAP[m]-3.0*AxBP[m]-2.0*AxAP[m])/6.0;
    PC[m+3][0][3][1] = ((6.0*BxBP[m]+6.0*BxAP[m])*y1*y1+(-6.0*BxBP[m]-3.0*BxAP[m])*y1)/6.0;
    PC[m+3][1][1][0] = ((-2.0*V0xBP[m]-2.0*BxV0[m]-BxAP[m]-2.0*AxV0[m]-AxBP[m]-AxAP[m])*y1
                        +2.0*V0xBP[m]+2.0*BxV0[m]+BxAP[m]+AxBP[m])/2.0;
    PC[m+3][1][1][1] = ((2.0*V0xBP[m]+AxBP[m])*y1)/2.0;
    PC[m+3][1][2][0] = -((-2.0*BxBP[m]-2.0*AxBP[m])*y1*y1
                         +(-2.0*V0xBP[m]-2.0*BxV0[m]+4*BxBP[m]-2.0*AxV0[m]+2.0*AxBP[m])*y1
                         +2.0*V0xBP[m]+V0xAP[m]+2.0*BxV0[m]-2.0*BxBP[m]+AxV0[m])/2.0;
    PC[m+3][1][2][1] = -((2.0*BxBP[m]+2.0*AxBP[m])*y1*y1+(2.0*V0xBP[m]-2.0*BxBP[m])*y1)/2.0;
    PC[m+3][1][3][0] = ((-6.0*BxBP[m]-6.0*AxBP[m])*y1*y1
                        +(12.0*BxBP[m]+3.0*BxAP[m]+9*AxBP[m]+3.0*AxAP[m])*y1-6.0*BxBP[m]
                        -3.0*BxAP[m]-3.0*AxBP[m]-2.0*AxAP[m])/6.0;
    PC[m+3][1][3][1] = ((6.0*BxBP[m]+6.0*AxBP[m])*y1*y1+(-6.0*BxBP[m]-3.0*AxBP[m])*y1)/6.0;
    PC[m+3][2][1][0] = -((2.0*BxV0[m]+BxAP[m])*y1-2.0*V0xBP[m]-2.0*BxV0[m]-BxAP[m]-AxBP[m])/2.0;
    PC[m+3][2][1][1] = -(2.0*V0xBP[m]+2.0*V0xAP[m]+AxBP[m]+AxAP[m])*y1/2.0;
    PC[m+3][2][2][0] = ((2.0*BxV0[m]-2.0*BxBP[m])*y1-2.0*V0xBP[m]-V0xAP[m]-2.0*BxV0[m]+2.0*BxBP[m]
                        -AxV0[m])/2.0;
    PC[m+3][2][2][1] = ((2.0*BxBP[m]+2.0*BxAP[m])*y1*y1
                        +(2.0*V0xBP[m]+2.0*V0xAP[m]-2.0*BxBP[m]-2.0*BxAP[m])*y1)/2.0;
    PC[m+3][2][3][0] = -((-6.0*BxBP[m]-3.0*BxAP[m])*y1+6.0*BxBP[m]+3.0*BxAP[m]+3.0*AxBP[m]
                         +2.0*AxAP[m])/6.0;
    PC[m+3][2][3][1] = -(((6.0*BxBP[m]+6.0*BxAP[m])*y1*y1
                          +(-6.0*BxBP[m]-6.0*BxAP[m]-3.0*AxBP[m]-3.0*AxAP[m])*y1))/6.0;
    PC[m+3][3][1][0] = -((2.0*V0xBP[m]+AxBP[m])*y1-2.0*V0xBP[m]-2.0*BxV0[m]-BxAP[m]-AxBP[m])/2.0;
    PC[m+3][3][1][1] = -(2.0*BxV0[m]+BxAP[m]+2.0*AxV0[m]+AxAP[m])*y1/2.0;
    PC[m+3][3][2][0] = ((2.0*V0xBP[m]-2.0*BxBP[m])*y1-2.0*V0xBP[m]-V0xAP[m]-2.0*BxV0[m]
                        +2.0*BxBP[m]-AxV0[m])/2.0;
    PC[m+3][3][2][1] = ((2.0*BxBP[m]+2.0*AxBP[m])*y1*y1
                        +(2.0*BxV0[m]-2.0*BxBP[m]+2.0*AxV0[m]-2.0*AxBP[m])*y1)/2.0;
    PC[m+3][3][3][0] = -((-6.0*BxBP[m]-3.0*AxBP[m])*y1+6.0*BxBP[m]+3.0*BxAP[m]+3.0*AxBP[m]
                         +2.0*AxAP[m])/6.0;
    PC[m+3][3][3][1] = -((6.0*BxBP[m]+6.0*AxBP[m])*y1*y1
                         +(-6.0*BxBP[m]-3.0*BxAP[m]-6.0*AxBP[m]-3.0*AxAP[m])*y1)/6.0;
    PC[m+3][4][1][0] = -((2.0*V0xBP[m]+AxBP[m])*y1-2.0*V0xBP[m]-2.0*BxV0[m]-BxAP[m]-AxBP[m])/2.0;
    PC[m+3][4][1][1] = -((2.0*V0xAP[m]+AxAP[m])*y1)/2.0;
    PC[m+3][4][2][0] = ((2.0*V0xBP[m]-2.0*BxBP[m])*y1-2.0*V0xBP[m]-V0xAP[m]-2.0*BxV0[m]
                        +2.0*BxBP[m]-AxV0[m])/2.0;
    PC[m+3][4][2][1] = (2.0*V0xAP[m]-2.0*BxAP[m])*y1/2.0;
    PC[m+3][4][3][0] = ((6.0*BxBP[m]+3.0*AxBP[m])*y1-6.0*BxBP[m]-3.0*BxAP[m]-3.0*AxBP[m]
                        -2.0*AxAP[m])/6.0;
    PC[m+3][4][3][1] = (6.0*BxAP[m]+3.0*AxAP[m])*y1/6.0;
    PC[m+3][5][1][0] = -((2.0*BxV0[m]+BxAP[m])*y1-2.0*V0xBP[m]-2.0*BxV0[m]-BxAP[m]-AxBP[m])/2.0;
    PC[m+3][5][1][1] = -(2.0*AxV0[m]+AxAP[m])*y1/2.0;
    PC[m+3][5][2][0] = ((2.0*BxV0[m]-2.0*BxBP[m])*y1-2.0*V0xBP[m]-V0xAP[m]-2.0*BxV0[m]
                        +2.0*BxBP[m]-AxV0[m])/2.0;
    PC[m+3][5][2][1] = (2.0*AxV0[m]-2.0*AxBP[m])*y1/2.0;
    PC[m+3][5][3][0] = ((6.0*BxBP[m]+3.0*BxAP[m])*y1-6.0*BxBP[m]-3.0*BxAP[m]-3.0*AxBP[m]
                        -2.0*AxAP[m])/6.0;
    PC[m+3][5][3][1] = (6.0*AxBP[m]+3.0*AxAP[m])*y1/6.0;
  }
  PX[0][0][0][0] = 1.0;  PX[0][0][1][0] = -1.0;
  PX[0][1][0][0] = 1.0;  PX[0][1][1][0] = -1.0;
  PX[0][2][0][0] = 1.0;  PX[0][2][1][0] = -1.0;
  PX[0][3][0][0] = 1.0;  PX[0][3][1][0] = -1.0;
  PX[0][4][0][0] = 1.0;  PX[0][4][1][0] = -1.0;
  PX[0][5][0][0] = 1.0;  PX[0][5][1][0] = -1.0;
  PX[1][0][0][0] = 1.0/2.0; PX[1][0][1][0] = -y1; PX[1][0][2][0] = (2.0*y1-1.0)/2.0;
  PX[1][1][0][0] = 1.0/2.0; PX[1][1][2][0] = -1.0/2.0;
  PX[1][2][0][0] = 1.0/2.0; PX[1][2][1][1] = -y1; PX[1][2][2][0] = -1.0/2.0; PX[1][2][2][1] = y1;
  PX[1][3][0][0] = 1.0/2.0; PX[1][3][2][0] = -1.0/2.0;
  PX[1][4][0][0] = 1.0/2.0; PX[1][4][1][1] = -y1; PX[1][4][2][0] = -1.0/2.0; PX[1][4][2][1] = y1;
  PX[1][5][0][0] = 1.0/2.0; PX[1][5][2][0] = -1.0/2.0;
  PX[2][0][1][0] = -y1+1.0; PX[2][0][2][0] = y1-1.0;
  PX[2][1][1][0] = -y1+1.0; PX[2][1][1][1] = y1; PX[2][1][2][0] = +y1-1.0; PX[2][1][2][1] = -y1;
  PX[2][2][1][0] = 1.0; PX[2][2][1][1] = -y1; PX[2][2][2][0] = -1.0; PX[2][2][2][1] = y1;
  PX[2][3][1][0] = -y1+1.0; PX[2][3][2][0] = y1-1.0;
  PX[2][4][1][0] = -y1+1.0; PX[2][4][2][0] = y1-1.0;
  PX[2][5][1][0] = 1.0; PX[2][5][2][0] = -1.0;
  PX[3][0][0][0] = 1.0/2.0; PX[3][0][2][0] = -1.0/2.0;
  PX[3][1][0][0] = 1.0/2.0; PX[3][1][1][0] = -y1; PX[3][1][2][0] = (2.0*y1-1.0)/2.0;
  PX[3][2][0][0] = 1.0/2.0; PX[3][2][2][0] = -1.0/2.0;
  PX[3][3][0][0] = 1.0/2.0; PX[3][3][1][1] = -y1; PX[3][3][2][0] = -1.0/2.0; PX[3][3][2][1] = y1;
  PX[3][4][0][0] = 1.0/2.0; PX[3][4][2][0] = -1.0/2.0;
  PX[3][5][0][0] = 1.0/2.0; PX[3][5][1][1] = -y1; PX[3][5][2][0] = -1.0/2.0; PX[3][5][2][1] = y1;
  PX[4][0][0][0] = 1.0/3.0; PX[4][0][1][0] = -y1/2.0; PX[4][0][3][0] = (3.0*y1-2.0)/6.0;
  PX[4][1][0][0] = 1.0/3.0; PX[4][1][1][0] = -y1/2.0; PX[4][1][3][0] = (3.0*y1-2.0)/6.0;
  PX[4][2][0][0] = 1.0/3.0; PX[4][2][1][1] = -y1/2.0; PX[4][2][3][0] = -1.0/3.0; PX[4][2][3][1] = y1/2.0;
  PX[4][3][0][0] = 1.0/3.0; PX[4][3][1][1] = -y1/2.0; PX[4][3][3][0] = -1.0/3.0; PX[4][3][3][1] = y1/2.0;
  PX[4][4][0][0] = 1.0/3.0; PX[4][4][1][1] = -y1/2.0; PX[4][4][3][0] = -1.0/3.0; PX[4][4][3][1] = y1/2.0;
  PX[4][5][0][0] = 1.0/3.0; PX[4][5][1][1] = -y1/2.0; PX[4][5][3][0] = -1.0/3.0; PX[4][5][3][1] = y1/2.0;
  PX[5][0][1][0] = -(y1-1.0)/2.0; PX[5][0][3][0] = (y1-1.0)/2.0;
  PX[5][1][1][0] = (-y1+1.0)/2.0; PX[5][1][1][1] = y1/2.0; PX[5][1][2][0] = +y1*y1-y1;
  PX[5][1][2][1] = -y1*y1; PX[5][1][3][0] = (-2.0*y1*y1+3.0*y1-1.0)/2.0; PX[5][1][3][1] = (2.0*y1*y1-y1)/2.0;
  PX[5][2][1][0] = -(-1.0)/2.0; PX[5][2][1][1] = -y1/2.0; PX[5][2][3][0] = (-1.0)/2.0; PX[5][2][3][1] = y1/2.0;
  PX[5][3][1][0] = -(y1-1.0)/2.0; PX[5][3][2][1] = (y1*y1-y1); PX[5][3][3][0] = -(-y1+1.0)/2.0;
  PX[5][3][3][1] = -(y1*y1-y1);
  PX[5][4][1][0] = -(y1-1.0)/2.0; PX[5][4][3][0] = (y1-1.0)/2.0;
  PX[5][5][1][0] = 1.0/2.0; PX[5][5][2][1] = -y1; PX[5][5][3][0] = -1.0/2.0; PX[5][5][3][1] = y1;
  PX[6][0][1][0] = -y1+1.0; PX[6][0][1][1] = y1; PX[6][0][2][0] = +y1-1.0; PX[6][0][2][1] = -y1;
  PX[6][1][1][0] = -y1+1.0; PX[6][1][2][0] = y1-1.0;
  PX[6][2][1][0] = -y1+1.0; PX[6][2][2][0] = y1-1.0;
  PX[6][3][1][0] = 1.0; PX[6][3][1][1] = -y1; PX[6][3][2][0] = -1.0; PX[6][3][2][1] = y1;
  PX[6][4][1][0] = 1.0; PX[6][4][2][0] = -1.0;
  PX[6][5][1][0] = -y1+1.0; PX[6][5][2][0] = y1-1.0;
  PX[7][0][1][0] = (-y1+1.0)/2.0; PX[7][0][1][1] = y1/2.0; PX[7][0][2][0] = y1*y1-y1;
  PX[7][0][2][1] = -y1*y1; PX[7][0][3][0] = (-2.0*y1*y1+3.0*y1-1.0)/2.0; PX[7][0][3][1] = (2.0*y1*y1-y1)/2.0;
  PX[7][1][1][0] = -(y1-1.0)/2.0; PX[7][1][3][0] = (y1-1.0)/2.0;
  PX[7][2][1][0] = -(y1-1.0)/2.0; PX[7][2][2][1] = (y1*y1-y1); PX[7][2][3][0] = -(-y1+1.0)/2.0;
  PX[7][2][3][1] = -(y1*y1-y1);
  PX[7][3][1][0] = -(-1.0)/2.0; PX[7][3][1][1] = -y1/2.0; PX[7][3][3][0] = -1.0/2.0; PX[7][3][3][1] = y1/2.0;
  PX[7][4][1][0] = 1.0/2.0; PX[7][4][2][1] = -y1; PX[7][4][3][0] = -1.0/2.0; PX[7][4][3][1] = y1;
  PX[7][5][1][0] = -(y1-1.0)/2.0; PX[7][5][3][0] = (y1-1.0)/2.0;
  PX[8][0][2][0] = +y1*y1-2.0*y1+1.0; PX[8][0][2][1] = (-y1*y1+y1);
  PX[8][0][3][0] = -y1*y1+2.0*y1-1.0; PX[8][0][3][1] = (y1*y1-y1);
  PX[8][1][2][0] = +y1*y1-2.0*y1+1.0; PX[8][1][2][1] = (-y1*y1+y1);
  PX[8][1][3][0] = -y1*y1+2.0*y1-1.0; PX[8][1][3][1] = (y1*y1-y1);
  PX[8][2][2][0] = -y1+1.0; PX[8][2][2][1] = (y1*y1-y1); PX[8][2][3][0] = y1-1.0; PX[8][2][3][1] = (-y1*y1+y1);
  PX[8][3][2][0] = -y1+1.0; PX[8][3][2][1] = (y1*y1-y1); PX[8][3][3][0] = +y1-1.0; PX[8][3][3][1] = (-y1*y1+y1);
  PX[8][4][2][0] = -y1+1.0; PX[8][4][3][0] = y1-1.0;
  PX[8][5][2][0] = -y1+1.0; PX[8][5][3][0] = y1-1.0;
}

__device__ void tdf_tables_cv(const FipW &W, double y1, double y2, double PC[6][6][4][2], double PX[9][6][4][2])
{
  const double *A = W.tA, *B = W.tB, *AP = W.tAP, *BP = W.tBP;
  for (int m = 0; m < 3; m++) {
    PC[m][0][1][0] = -AP[m] * y2 + B[m] * y1 + A[m];
    PC[m][0][1][1] = -BP[m] * y2;
    PC[m][0][2][0] = 0.0; PC[m][0][2][1] = 0.0;
    PC[m][1][1][0] = A[m] * y2 - BP[m] * y1 - AP[m];
    PC[m][1][1][1] = B[m] * y2;
    PC[m][1][2][0] = 0.0; PC[m][1][2][1] = 0.0;
    PC[m + 3][0][1][0] = W.V0xAP[m] * y2 + W.BxV0[m] * y1 + W.AxV0[m];
    PC[m + 3][0][1][1] = W.V0xBP[m] * y2;
    PC[m + 3][0][2][0] = (W.BxAP[m] * y1 + W.AxAP[m]) * y2;
    PC[m + 3][0][2][1] = (W.BxBP[m] * y1 + W.AxBP[m]) * y2;
    PC[m + 3][1][1][0] = W.AxV0[m] * y2 + W.V0xBP[m] * y1 + W.V0xAP[m];
    PC[m + 3][1][1][1] = W.BxV0[m] * y2;
    PC[m + 3][1][2][0] = (W.AxBP[m] * y1 + W.AxAP[m]) * y2;
    PC[m + 3][1][2][1] = (W.BxBP[m] * y1 + W.BxAP[m]) * y2;
  }
  PX[0][0][0][0] = 1.0;   PX[0][1][0][0] = 1.0;
  PX[1][0][1][0] = y2;    PX[1][1][1][0] = 1.0;
  PX[2][0][1][1] = y2;    PX[2][1][1][0] = y1;
  PX[3][0][1][0] = 1.0;   PX[3][1][1][0] = y2;
  PX[4][0][2][0] = y2;    PX[4][1][2][0] = y2;
  PX[5][0][2][1] = y2;    PX[5][1][2][0] = y1 * y2;
  PX[6][0][1][0] = y1;    PX[6][1][1][1] = y2;
  PX[7][0][2][0] = y1 * y2; PX[7][1][2][1] = y2;
  PX[8][0][2][1] = y1 * y2; PX[8][1][2][1] = y1 * y2;
}

__device__ void tdf_sum(const FipW &W, const double *y, double *f)
{
  int NumRegions, nOffset, nMinC, nMaxC, nMinX, nMaxX;
  double Jac;
  if (W.ce) { NumRegions = 6; nOffset = 2; Jac = y[0]; nMinC = 1; nMaxC = 3; nMinX = 0; nMaxX = 3; }
  else { NumRegions = 2; nOffset = 3; Jac = y[1]; nMinC = 1; nMaxC = 2; nMinX = 0; nMaxX = 2; }
  double A[6], B[6], G2[6];
  tdf_abg(W, y, A, B, G2);
  double JrM3[6][7], LrM3[6][7], Jrp[3][6][7], Lrp[3][6][7];
  double Iq[4], Iy[4];
  for (int d = 0; d < NumRegions; d++) {
    tdf_qfp(B[d], G2[d], Iq, Iy);
    double A2 = A[d] * A[d], A3 = A2 * A[d];
    for (int n = nMinC + nOffset; n <= nMaxC + nOffset; n++) {
      JrM3[d][n] = Iq[0] / (A3 * (1.0 + n - 3.0));
      LrM3[d][n] = Iy[0] / (A3 * (1.0 + n - 3.0));
    }
    for (int n = nMinX + nOffset; n <= nMaxX + nOffset; n++) {
      Jrp[0][d][n] = Iq[1] / (A[d] * (1.0 + n - 1.0));
      Lrp[0][d][n] = Iy[1] / (A[d] * (1.0 + n - 1.0));
      Jrp[1][d][n] = A[d] * Iq[2] / (1.0 + n + 1.0);
      Lrp[1][d][n] = A[d] * Iy[2] / (1.0 + n + 1.0);
      Jrp[2][d][n] = A2 * Iq[3] / (1.0 + n + 2.0);
      Lrp[2][d][n] = A2 * Iy[3] / (1.0 + n + 2.0);
    }
  }
  double PC[6][6][4][2], PX[9][6][4][2];
  for (int a = 0; a < 6; a++) for (int b = 0; b < 6; b++) for (int c = 0; c < 4; c++) { PC[a][b][c][0] = 0; PC[a][b][c][1] = 0; }
  for (int a = 0; a < 9; a++) for (int b = 0; b < 6; b++) for (int c = 0; c < 4; c++) { PX[a][b][c][0] = 0; PX[a][b][c][1] = 0; }
  if (W.ce) tdf_tables_ce(W, y[0], PC, PX); else tdf_tables_cv(W, y[0], y[1], PC, PX);
  for (int i = 0; i < NF; i++) f[i] = 0.0;
  int ns = 0;
  for (int np = 0; np < 6; np++, ns++)
    for (int d = 0; d < NumRegions; d++)
      for (int n = nMinC; n <= nMaxC; n++)
        f[ns] += Jac * (PC[np][d][n][0] * JrM3[d][n + nOffset] + PC[np][d][n][1] * LrM3[d][n + nOffset]);
  for (int nk = 0; nk < 3; nk++)
    for (int np = 0; np < 9; np++, ns++)
      for (int d = 0; d < NumRegions; d++)
        for (int n = nMinX; n <= nMaxX; n++)
          f[ns] += Jac * (PX[np][d][n][0] * Jrp[nk][d][n + nOffset] + PX[np][d][n][1] * Lrp[nk][d][n + nOffset]);
}

// ------------------------- rule geometry ------------------------------
// coordinates of point p of the rule on region (c, h); dim 1: GK15,
// dims 2, 3: Genz-Malik in the order of hcubature.c:385-448
__device__ void rule_point(int dim, int p, const double *c, const double *h, double *x)
{
  const double lambda2 = 0.3585685828003180919906451539079374954541;
  const double lambda4 = 0.9486832980505137995996680633298155601160;
  const double lambda5 = 0.6882472016116852977216287342936235251269;
  if (dim == 1) {
    const double xgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                           0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                           0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
                           0.207784955007898467600689403773245, 0.0};
    if (p == 0) { x[0] = c[0]; return; }
    int q = p - 1;
    int j2 = (q < 6) ? 2 * (q / 2) + 1 : 2 * ((q - 6) / 2);
    double w = h[0] * xgk[j2];
    x[0] = (q & 1) ? c[0] + w : c[0] - w;
    return;
  }
  for (int i = 0; i < dim; i++) x[i] = c[i];
  if (p == 0) return;
  int q = p - 1;
  if (q < 4 * dim) {
    int i = q / 4, s = q % 4;
    double l = (s < 2) ? h[i] * lambda2 : h[i] * lambda4;
    x[i] = (s & 1) ? c[i] + l : c[i] - l;
    return;
  }
  q -= 4 * dim;
  int nRR = 2 * dim * (dim - 1);
  if (q < nRR) {
    int pairIdx = q / 4, s = q % 4, cnt = 0;
    for (int i = 0; i + 1 < dim; i++)
      for (int j = i + 1; j < dim; j++, cnt++)
        if (cnt == pairIdx) {
          double li = h[i] * lambda4, lj = h[j] * lambda4;
          // (-i,-j), (+i,-j), (+i,+j), (-i,+j)
          x[i] = (s == 1 || s == 2) ? c[i] + li : c[i] - li;
          x[j] = (s >= 2) ? c[j] + lj : c[j] - lj;
        }
    return;
  }
  q -= nRR;
  unsigned g = (unsigned)q ^ ((unsigned)q >> 1);
  for (int i = 0; i < dim; i++) {
    double l = h[i] * lambda5;
    x[i] = ((g >> i) & 1) ? c[i] - l : c[i] + l;
  }
}

__device__ inline void integrand(const FipW &W, const double *x, double *f)
{
  if (W.ncv == 0) cfd3(W, x, f); else tdf_sum(W, x, f);
}

// per-function rule sums for one region whose point values are
// buf[p*NF + j]; returns value, error and the per-dimension
// fourth-difference contribution of function j (hcubature.c:508-560)
__device__ void rule_sums(int dim, int j, const double *buf, double vol, double *val, double *err, double *diff)
{
  if (dim == 1) {
    const unsigned n = 8;
    const double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                          0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
    const double wgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                           0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                           0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                           0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
    const double hw = vol * 0.5;   // halfwidth (vol = 2 h in 1-D)
#define VK(q) buf[(q) * NF + j]
    double rg = VK(0) * wg[n / 2 - 1], rk = VK(0) * wgk[n - 1], rabs = fabs(rk), rasc, mean, e;
    unsigned q = 1, jj;
    for (jj = 0; jj < (n - 1) / 2; jj++) {
      int j2 = 2 * jj + 1;
      double v = VK(q) + VK(q + 1);
      rg += wg[jj] * v; rk += wgk[j2] * v;
      rabs += wgk[j2] * (fabs(VK(q)) + fabs(VK(q + 1)));
      q += 2;
    }
    for (jj = 0; jj < n / 2; jj++) {
      int j2 = 2 * jj;
      rk += wgk[j2] * (VK(q) + VK(q + 1));
      rabs += wgk[j2] * (fabs(VK(q)) + fabs(VK(q + 1)));
      q += 2;
    }
    *val = rk * hw;
    mean = rk * 0.5;
    rasc = wgk[n - 1] * fabs(VK(0) - mean);
    q = 1;
    for (jj = 0; jj < (n - 1) / 2; jj++) {
      int j2 = 2 * jj + 1;
      rasc += wgk[j2] * (fabs(VK(q) - mean) + fabs(VK(q + 1) - mean));
      q += 2;
    }
    for (jj = 0; jj < n / 2; jj++) {
      int j2 = 2 * jj;
      rasc += wgk[j2] * (fabs(VK(q) - mean) + fabs(VK(q + 1) - mean));
      q += 2;
    }
#undef VK
    e = fabs(rk - rg) * hw;
    rabs *= hw; rasc *= hw;
    if (rasc != 0 && e != 0) {
      double scale = pow((200 * e / rasc), 1.5);
      e = (scale < 1) ? rasc * scale : rasc;
    }
    if (rabs > 2.2250738585072014e-308 / (50 * 2.220446049250313e-16)) {
      double min_err = 50 * 2.220446049250313e-16 * rabs;
      if (min_err > e) e = min_err;
    }
    *err = e;
    diff[0] = 0.0;
    return;
  }
  const double lambda2 = 0.3585685828003180919906451539079374954541;
  const double lambda4 = 0.9486832980505137995996680633298155601160;
  const double weight2 = 980. / 6561., weight4 = 200. / 19683.;
  const double weightE2 = 245. / 486., weightE4 = 25. / 729.;
  const double ratio = (lambda2 * lambda2) / (lambda4 * lambda4);
  double weight1 = (double)(12824 - 9120 * dim + 400 * dim * dim) / 19683.;
  double weight3 = (double)(1820 - 400 * dim) / 19683.;
  double weight5 = 6859. / 19683. / (double)(1U << dim);
  double weightE1 = (double)(729 - 950 * dim + 50 * dim * dim) / 729.;
  double weightE3 = (double)(265 - 100 * dim) / 1458.;
#define VALS(i) buf[(i) * NF + j]
  double val0 = VALS(0), sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
  int k, k0 = 1;
  for (k = 0; k < dim; k++) {
    double v0 = VALS(k0 + 4 * k), v1 = VALS(k0 + 4 * k + 1), v2 = VALS(k0 + 4 * k + 2), v3 = VALS(k0 + 4 * k + 3);
    sum2 += v0 + v1;
    sum3 += v2 + v3;
    diff[k] = fabs(v0 + v1 - 2 * val0 - ratio * (v2 + v3 - 2 * val0));
  }
  k0 += 4 * k;
  int nRR = 2 * dim * (dim - 1), n5 = 1 << dim;
  for (k = 0; k < nRR; k++) sum4 += VALS(k0 + k);
  k0 += k;
  for (k = 0; k < n5; k++) sum5 += VALS(k0 + k);
#undef VALS
  double result = vol * (weight1 * val0 + weight2 * sum2 + weight3 * sum3 + weight4 * sum4 + weight5 * sum5);
  double res5th = vol * (weightE1 * val0 + weightE2 * sum2 + weightE3 * sum3 + weightE4 * sum4);
  *val = result;
  *err = fabs(res5th - result);
}

// ------------------------------ heap ----------------------------------
struct HeapRef { double *key; int *slot; };

__device__ void heap_push(HeapRef h, int &n, double key, int slot)
{
  int ins = n++;
  while (ins) {
    int parent = (ins - 1) / 2;
    if (key <= h.key[parent]) break;
    h.key[ins] = h.key[parent];
    h.slot[ins] = h.slot[parent];
    ins = parent;
  }
  h.key[ins] = key;
  h.slot[ins] = slot;
}

__device__ int heap_pop(HeapRef h, int &n)
{
  int ret = h.slot[0];
  int i = 0, nn = --n, child;
  h.key[0] = h.key[nn];
  h.slot[0] = h.slot[nn];
  while ((child = i * 2 + 1) < nn) {
    int largest;
    if (h.key[child] <= h.key[i]) largest = i; else largest = child;
    if (++child < nn && h.key[largest] < h.key[child]) largest = child;
    if (largest == i) break;
    double tk = h.key[i]; int ts = h.slot[i];
    h.key[i] = h.key[largest]; h.slot[i] = h.slot[largest];
    h.key[largest] = tk; h.slot[largest] = ts;
    i = largest;
  }
  return ret;
}

__device__ void setup_w(const FippiJob &J, FipW &W)
{
  W.ncv = J.ncv;
  const double *a0 = J.OVa, *a1 = J.OVa + 3, *a2 = J.OVa + 6;
  const double *b0 = J.OVb, *b1 = J.OVb + 3, *b2 = J.OVb + 6;
  if (J.ncv == 0) {
    for (int m = 0; m < 3; m++) {
      W.V0[m] = a0[m]; W.A[m] = a1[m] - a0[m]; W.B[m] = a2[m] - a1[m];
      W.V0P[m] = b0[m]; W.AP[m] = b1[m] - b0[m]; W.BP[m] = b2[m] - b1[m];
    }
    W.ce = 0;
    return;
  }
  // V1=OVa0, V2=OVa1, V3=OVa2, V2P=OVb1, V3P=OVb2
  const double *V1 = a0, *V2 = a1, *V3 = a2, *V2P = b1, *V3P = b2;
  double L[3];
  sub3(V2, V1, W.tA); sub3(V3, V2, W.tB); sub3(V2P, V1, W.tAP); sub3(V3P, V2P, W.tBP);
  sub3(W.tBP, W.tB, L);
  W.A2 = dot3(W.tA, W.tA); W.B2 = dot3(W.tB, W.tB); W.AP2 = dot3(W.tAP, W.tAP); W.BP2 = dot3(W.tBP, W.tBP);
  W.L2 = dot3(L, L); W.AdB = dot3(W.tA, W.tB); W.AdAP = dot3(W.tA, W.tAP); W.AdBP = dot3(W.tA, W.tBP);
  W.AdL = dot3(W.tA, L); W.BdAP = dot3(W.tB, W.tAP); W.BdBP = dot3(W.tB, W.tBP); W.APdBP = dot3(W.tAP, W.tBP);
  W.BPdL = dot3(W.tBP, L);
  cross3(V1, W.tAP, W.V0xAP); cross3(V1, W.tBP, W.V0xBP);
  cross3(W.tA, V1, W.AxV0); cross3(W.tB, V1, W.BxV0);
  cross3(W.tA, W.tAP, W.AxAP); cross3(W.tB, W.tAP, W.BxAP);
  cross3(W.tA, W.tBP, W.AxBP); cross3(W.tB, W.tBP, W.BxBP);
  W.ce = ((float)V2[0] == (float)V2P[0]) && ((float)V2[1] == (float)V2P[1]) && ((float)V2[2] == (float)V2P[2]);
}

// ------------------------------ kernel --------------------------------
// one 64-lane block per wave slot; jobs pulled from a counter
__global__ void __launch_bounds__(64)
k_fippi(const FippiJob *jobs, int njobs, const int *jobList, double *records,
        double *heapData, int *heapSlot, double *heapKey, int cap, int *counter, int *overflow)
{
  __shared__ double buf[66 * NF];
  __shared__ double dpart[2][NF][3];
  __shared__ double sreg[2][8];     // center[3] half[3] vol split for A, B
  __shared__ double eA[2][NF], rA[2][NF];
  __shared__ int s_job, s_pop, s_slotB;
  __shared__ int s_stop;
  // the job's constants, the same for every lane: in LDS, not in each
  // lane's registers
  __shared__ FipW sW;
  const int lane = threadIdx.x;
  double *data = heapData + (long long)blockIdx.x * cap * REC;
  HeapRef H;
  H.key = heapKey + (long long)blockIdx.x * cap;
  H.slot = heapSlot + (long long)blockIdx.x * cap;

  for (;;) {
    if (lane == 0) s_job = atomicAdd(counter, 1);
    __syncthreads();
    int jj = s_job;
    if (jj >= njobs) break;
    int job = jobList ? jobList[jj] : jj;
    const FippiJob &J = jobs[job];
    if (lane == 0) setup_w(J, sW);
    __syncthreads();
    const FipW &W = sW;
    const int dim = (J.ncv == 0) ? 3 : (W.ce ? 1 : 2);
    const int npts = (dim == 1) ? 15 : 1 + 4 * dim + 2 * dim * (dim - 1) + (1 << dim);
    const long maxEval = (J.ncv == 0) ? 1000000 : 10000;
    const double reqAbs = (J.ncv == 0) ? 1.0e-10 : 0.0, reqRel = (J.ncv == 0) ? 1.0e-6 : 1.0e-5;
    // running totals held by lanes 0..32
    double totV = 0.0, totE = 0.0;
    int n = 0, nused = 0;
    long numEval = 0;
    bool ovf = false;
    // initial region
    if (lane == 0) {
      for (int i = 0; i < 3; i++) { sreg[0][i] = 0.5; sreg[0][3 + i] = 0.5; }
      double v = 1.0;
      for (int i = 0; i < dim; i++) v *= 2 * 0.5;
      sreg[0][6] = v; sreg[0][7] = 0;
    }
    __syncthreads();
    for (int p = lane; p < npts; p += 64) {
      double x[3];
      rule_point(dim, p, sreg[0], sreg[0] + 3, x);
      double f[NF];
      integrand(W, x, f);
      for (int q = 0; q < NF; q++) buf[p * NF + q] = f[q];
    }
    __syncthreads();
    if (lane < NF) {
      double v, e, d[3] = {0, 0, 0};
      rule_sums(dim, lane, buf, sreg[0][6], &v, &e, d);
      eA[0][lane] = v; rA[0][lane] = e;
      for (int q = 0; q < 3; q++) dpart[0][lane][q] = d[q];
    }
    __syncthreads();
    if (lane == 0) {
      double dsum[3] = {0, 0, 0}, em = 0;
      for (int q = 0; q < NF; q++) { for (int i = 0; i < dim; i++) dsum[i] += dpart[0][q][i]; if (rA[0][q] > em) em = rA[0][q]; }
      double mx = 0; int sd = 0;
      for (int i = 0; i < dim; i++) if (dsum[i] > mx) { mx = dsum[i]; sd = i; }
      if (dim == 1) sd = 0;
      double *R = data;
      for (int i = 0; i < 6; i++) R[i] = sreg[0][i];
      R[6] = sreg[0][6]; R[7] = sd;
      heap_push(H, n, em, 0);
    }
    if (lane < NF) { data[8 + lane] = eA[0][lane]; data[8 + NF + lane] = rA[0][lane]; totV += eA[0][lane]; totE += rA[0][lane]; }
    nused = 1;
    numEval = npts;
    __syncthreads();

    while (numEval < maxEval) {
      // convergence (converged.h ERROR_INDIVIDUAL)
      int bad = (lane < NF) && (totE > reqAbs && totE > fabs(totV) * reqRel);
      if (!__any(bad)) break;
      if (nused + 1 > cap) { ovf = true; break; }
      if (lane == 0) { s_pop = heap_pop(H, n); s_slotB = nused; }
      __syncthreads();
      int sa = s_pop, sb = s_slotB;
      nused++;
      double *RA = data + (long long)sa * REC, *RB = data + (long long)sb * REC;
      if (lane < NF) { totV -= RA[8 + lane]; totE -= RA[8 + NF + lane]; }
      if (lane == 0) {
        // cut_region (hcubature.c:215-228)
        int d = (int)RA[7];
        double c[3] = {RA[0], RA[1], RA[2]}, h[3] = {RA[3], RA[4], RA[5]};
        double vol = RA[6];
        h[d] *= 0.5;
        vol *= 0.5;
        double volB = 1;
        for (int i = 0; i < dim; i++) volB *= 2 * h[i];
        double cA[3] = {c[0], c[1], c[2]}, cB[3] = {c[0], c[1], c[2]};
        cA[d] -= h[d];
        cB[d] += h[d];
        for (int i = 0; i < 3; i++) { sreg[0][i] = cA[i]; sreg[0][3 + i] = h[i]; sreg[1][i] = cB[i]; sreg[1][3 + i] = h[i]; }
        sreg[0][6] = vol; sreg[1][6] = volB;
      }
      __syncthreads();
      for (int g = lane; g < 2 * npts; g += 64) {
        int rg = g / npts, p = g % npts;
        double x[3];
        rule_point(dim, p, sreg[rg], sreg[rg] + 3, x);
        double f[NF];
        integrand(W, x, f);
        for (int q = 0; q < NF; q++) buf[g * NF + q] = f[q];
      }
      __syncthreads();
      if (lane < NF) {
        for (int rg = 0; rg < 2; rg++) {
          double v, e, d[3] = {0, 0, 0};
          rule_sums(dim, lane, buf + rg * npts * NF, sreg[rg][6], &v, &e, d);
          eA[rg][lane] = v; rA[rg][lane] = e;
          for (int q = 0; q < 3; q++) dpart[rg][lane][q] = d[q];
        }
      }
      __syncthreads();
      if (lane < NF) {
        RA[8 + lane] = eA[0][lane]; RA[8 + NF + lane] = rA[0][lane];
        RB[8 + lane] = eA[1][lane]; RB[8 + NF + lane] = rA[1][lane];
        totV += eA[0][lane]; totE += rA[0][lane];
        totV += eA[1][lane]; totE += rA[1][lane];
      }
      if (lane == 0) {
        double em[2];
        int sdim[2];
        for (int rg = 0; rg < 2; rg++) {
          double dsum[3] = {0, 0, 0}, m = 0;
          for (int q = 0; q < NF; q++) { for (int i = 0; i < dim; i++) dsum[i] += dpart[rg][q][i]; if (rA[rg][q] > m) m = rA[rg][q]; }
          double mx = 0; int sd = 0;
          for (int i = 0; i < dim; i++) if (dsum[i] > mx) { mx = dsum[i]; sd = i; }
          em[rg] = m; sdim[rg] = sd;
        }
        for (int i = 0; i < 6; i++) { RA[i] = sreg[0][i]; RB[i] = sreg[1][i]; }
        RA[6] = sreg[0][6]; RB[6] = sreg[1][6];
        RA[7] = sdim[0]; RB[7] = sdim[1];
        heap_push(H, n, em[0], sa);
        heap_push(H, n, em[1], sb);
      }
      numEval += 2 * npts;
      __syncthreads();
    }
    // broadcast the heap size held by lane 0
    if (lane == 0) s_stop = n;
    __syncthreads();
    int nh = s_stop;
    if (ovf) {
      if (lane == 0) overflow[job] = 1;
    } else if (lane < NF) {
      double v = 0;
      for (int i = 0; i < nh; i++) v += data[(long long)H.slot[i] * REC + 8 + lane];
      double *out = records + (long long)job * SCUFFX_QIF_LEN;
      out[lane] = v;
    }
    if (!ovf && lane < 3) {
      double *out = records + (long long)job * SCUFFX_QIF_LEN;
      out[33 + lane] = (J.OVa[lane] + J.OVa[3 + lane] + J.OVa[6 + lane]) / 3.0;
    }
    __syncthreads();
  }
}

hipError_t launch_fippi(hipStream_t stream, const FippiJob *jobs, int njobs, double *records,
                        double *heapData, int *heapSlot, double *heapKey, int cap, int nslots,
                        int *counter, int *overflow, const int *jobList)
{
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fippi, dim3(nslots), dim3(64), 0, stream, jobs, njobs, jobList, records,
                     heapData, heapSlot, heapKey, cap, counter, overflow);
  return hipGetLastError();
}

}  // namespace scuffx
