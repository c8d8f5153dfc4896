/*
 * td.c -- oracle: generalized Taylor-Duffy integration of singular
 * panel-panel integrals.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates libs/libscuff/TaylorDuffy.cc:152-367 (integrand and driver),
 * :373-504 (alpha/beta/gamma), :506-566 (X), :568-634 (script P),
 * :636-778 (quadratic-form integrals), :780-838 (script J/L),
 * :840-935 (ExpRelV3P0, script K), :940-1262 (Upsilon tables),
 * :1265-1528 (geometric parameters, P = UNITY / PMCHWG1 / PMCHWC).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NREG 6
#define NWP 5
#define NMON 22
#define NP 4   /* P polynomials restated: 0 UNITY, 2 PMCHWG1, 3 PMCHWC */

/* monomial slots (TaylorDuffy.cc:52-75) */
enum { T1 = 0, TY1 = 1, TY12 = 2, TY2 = 3, TY1Y2 = 4, TY22 = 5, TY13 = 6,
       TY12Y2 = 7, TY1Y22 = 8, TY23 = 9, TY12Y22 = 10, TY13Y2 = 11, TY13Y22 = 12,
       TY3 = 13, TY1Y3 = 14, TY2Y3 = 15, TY12Y3 = 16, TY1Y2Y3 = 17, TY22Y3 = 18,
       TY12Y2Y3 = 19, TY1Y22Y32 = 20, TY22Y32 = 21 };

typedef struct {
  int WhichCase, NumPKs;
  const int *PIndex, *KIndex;
  const zdouble *KParam;
  int TwiceIntegrable;
  double A2, B2, AP2, BP2, L2, AdB, AdAP, AdBP, AdL, BdAP, BdBP, APdBP, BPdL;
  double U[NP][NREG][NWP][NMON];
  int nMin[NP], nMax[NP];
  int NeedP[NP];
  int nCalls;
} tdw;

static double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static void sub3(const double *a, const double *b, double *c) { c[0]=a[0]-b[0]; c[1]=a[1]-b[1]; c[2]=a[2]-b[2]; }
static double *cross3(const double *a, const double *b, double *c)
{ c[0]=a[1]*b[2]-a[2]*b[1]; c[1]=a[2]*b[0]-a[0]*b[2]; c[2]=a[0]*b[1]-a[1]*b[0]; return c; }

/* Px^2 + 2Qx + R -> A^2[(x+B)^2 + G^2]  (TaylorDuffy.cc:373-387) */
static void pqr_abg2(double P, double Q, double R, double *A, double *B, double *G2)
{
  *A = sqrt(P);
  *B = Q/P;
  double RoP = R/P;
  *G2 = RoP - (*B)*(*B);
  if (fabs(*G2) < 1.0e-6*RoP) { *B = 0.0; *G2 = RoP; }
}

static void get_abg2(const tdw *W, const double *y, double *A, double *B, double *G2)
{
  if (W->WhichCase == OTD_COMMONTRIANGLE) {
    double A2 = W->A2, AdB = W->AdB, B2 = W->B2;
    double C2 = A2 + B2 + 2.0*AdB, CdB = AdB + B2;
    A[0] = sqrt(B2); A[1] = sqrt(C2); A[2] = sqrt(A2);
    B[0] = AdB/B2; B[1] = -CdB/C2; B[2] = AdB/A2;
    G2[0] = A2/B2 - (AdB*AdB)/(B2*B2);
    G2[1] = B2/C2 - (CdB*CdB)/(C2*C2);
    G2[2] = B2/A2 - (AdB*AdB)/(A2*A2);
  } else if (W->WhichCase == OTD_COMMONEDGE) {
    double y1 = y[0], y12 = y1*y1;
    double A2 = W->A2, BP2 = W->BP2, L2 = W->L2, AdBP = W->AdBP, AdL = W->AdL, BPdL = W->BPdL;
    pqr_abg2((BP2 - 2.0*BPdL + L2)*y12,
             y1*(L2 + BPdL*(y1-1.0) - (AdL + L2 - AdBP)*y1),
             L2 - 2*AdL*y1 - 2*L2*y1 + A2*y12 + 2*AdL*y12 + L2*y12, A+0, B+0, G2+0);
    pqr_abg2(BP2*y12,
             (BPdL + AdBP*y1 - BPdL*y1)*y1,
             L2 + 2*AdL*y1 - 2*L2*y1 + (A2 - 2*AdL + L2)*y12, A+1, B+1, G2+1);
    pqr_abg2((A2 + 2*AdBP + BP2)*y12,
             y1*(AdL*(-1 + y1) + BPdL*(-1 + y1) - (AdBP + BP2)*y1),
             L2 + 2*BPdL*y1 - 2*L2*y1 + BP2*y12 - 2*BPdL*y12 + L2*y12, A+2, B+2, G2+2);
    pqr_abg2((A2 + 2*AdBP - 2*AdL + BP2 - 2*BPdL + L2)*y12,
             y1*(AdL - L2 - (AdBP + BP2)*y1 + BPdL*(1 + y1)),
             L2 - 2*BPdL*y1 + BP2*y12, A+3, B+3, G2+3);
    pqr_abg2(A2*y12,
             (AdBP*y1 - AdL)*y1,
             L2 - 2*BPdL*y1 + BP2*y12, A+4, B+4, G2+4);
    pqr_abg2(A2*y12,
             y1*(AdL + AdBP*y1 - AdL*y1),
             L2 + 2*BPdL*y1 - 2*L2*y1 + BP2*y12 - 2*BPdL*y12 + L2*y12, A+5, B+5, G2+5);
  } else {
    double y1 = y[0], y2 = y[1], y22 = y2*y2;
    double A2 = W->A2, B2 = W->B2, AP2 = W->AP2, BP2 = W->BP2, AdB = W->AdB, AdAP = W->AdAP,
           AdBP = W->AdBP, BdAP = W->BdAP, BdBP = W->BdBP, APdBP = W->APdBP;
    pqr_abg2(BP2*y22,
             -AdBP*y2 - BdBP*y1*y2 + APdBP*y22,
             A2 + B2*y1*y1 + AP2*y2*y2 + 2.0*AdB*y1 - 2.0*AdAP*y2 - 2.0*BdAP*y1*y2, A+0, B+0, G2+0);
    pqr_abg2(B2*y22,
             AdB*y22 - BdAP*y2 - BdBP*y1*y2,
             A2*y2*y2 + AP2 + BP2*y1*y1 - 2*AdAP*y2 - 2*AdBP*y2*y1 + 2*APdBP*y1, A+1, B+1, G2+1);
  }
}

/* distance function X_d(y) for once-integrable kernels */
static double ce_x(const tdw *W, double u1, double u2, double xi2)
{
  return sqrt(u1*u1*W->A2 + u2*u2*W->BP2 + xi2*xi2*W->L2
              + 2.0*(u1*(u2*W->AdBP + xi2*W->AdL) + u2*xi2*W->BPdL));
}
static double cv_x(const tdw *W, double xi1, double xi2, double eta1, double eta2)
{
  return sqrt(xi1*xi1*W->A2 + xi2*xi2*W->B2 + eta1*eta1*W->AP2 + eta2*eta2*W->BP2
              + 2.0*xi1*(xi2*W->AdB - eta1*W->AdAP - eta2*W->AdBP)
              - 2.0*xi2*(eta1*W->BdAP + eta2*W->BdBP)
              + 2.0*eta1*eta2*W->APdBP);
}
static void get_x(const tdw *W, const double *y, double *X)
{
  if (W->WhichCase == OTD_COMMONTRIANGLE) {
    double t = y[0], u1, u2;
    u1 = 1.0; u2 = t;       X[0] = sqrt(W->A2*u1*u1 + 2.0*W->AdB*u1*u2 + W->B2*u2*u2);
    u1 = t;   u2 = (t-1);   X[1] = sqrt(W->A2*u1*u1 + 2.0*W->AdB*u1*u2 + W->B2*u2*u2);
    u1 = t;   u2 = 1.0;     X[2] = sqrt(W->A2*u1*u1 + 2.0*W->AdB*u1*u2 + W->B2*u2*u2);
  } else if (W->WhichCase == OTD_COMMONEDGE) {
    double y1 = y[0], y2 = y[1];
    X[0] = ce_x(W, -y1,     -y1*y2,         1.0 - y1 + y1*y2);
    X[1] = ce_x(W,  y1,      y1*y2,         1.0 - y1);
    X[2] = ce_x(W, -y1*y2,   y1*(1.0 - y2), 1.0 - y1);
    X[3] = ce_x(W,  y1*y2,  -y1*(1.0 - y2), 1.0 - y1*y2);
    X[4] = ce_x(W, -y1*y2,  -y1,            1.0);
    X[5] = ce_x(W,  y1*y2,   y1,            1.0 - y1);
  } else {
    double y1 = y[0], y2 = y[1], y3 = y[2];
    X[0] = cv_x(W, 1.0, y1, y2, y2*y3);
    X[1] = cv_x(W, y2, y2*y3, 1.0, y1);
  }
}

static void get_script_p(const tdw *W, int np, const double *y, double P[NREG][NWP][3])
{
  int nMin = W->nMin[np], nMax = W->nMax[np];
  const double (*U)[NWP][NMON] = W->U[np];
  if (W->WhichCase == OTD_COMMONTRIANGLE) {
    for (int d = 0; d < 3; d++)
      for (int n = nMin; n <= nMax; n++) {
        P[d][n][0] = U[d][n][T1]; P[d][n][1] = U[d][n][TY1]; P[d][n][2] = U[d][n][TY12];
      }
  } else if (W->WhichCase == OTD_COMMONEDGE) {
    double y1 = y[0], y12 = y1*y1;
    for (int d = 0; d < 6; d++)
      for (int n = nMin; n <= nMax; n++) {
        P[d][n][0] = U[d][n][T1] + U[d][n][TY1]*y1 + U[d][n][TY12]*y12;
        P[d][n][1] = U[d][n][TY2] + U[d][n][TY1Y2]*y1 + U[d][n][TY12Y2]*y12;
        P[d][n][2] = U[d][n][TY22] + U[d][n][TY1Y22]*y1 + U[d][n][TY12Y22]*y12;
      }
  } else {
    double y1 = y[0], y12 = y1*y1, y2 = y[1], y22 = y2*y2;
    for (int d = 0; d < 2; d++)
      for (int n = nMin; n <= nMax; n++) {
        P[d][n][0] = U[d][n][T1] + U[d][n][TY1]*y1 + U[d][n][TY2]*y2 + U[d][n][TY12]*y12
                   + U[d][n][TY1Y2]*y1*y2 + U[d][n][TY22]*y22 + U[d][n][TY12Y2]*y12*y2
                   + U[d][n][TY1Y22]*y1*y22;
        P[d][n][1] = U[d][n][TY2Y3]*y2 + U[d][n][TY1Y2Y3]*y1*y2 + U[d][n][TY22Y3]*y22
                   + U[d][n][TY12Y2Y3]*y12*y2;
        P[d][n][2] = U[d][n][TY22Y32]*y22 + U[d][n][TY1Y22Y32]*y1*y22;
      }
  }
}

/* \int_0^1 [(y+P)^2+Q2]^{p/2} dy and \int_0^1 y[...]^{p/2} dy
   (TaylorDuffy.cc:636-778) */
static void qfp_integral(double P, double Q2, int p, double *I0, double *I1)
{
  double P2 = P*P, PP1 = P + 1.0, PP12 = PP1*PP1, Q4 = Q2*Q2;
  double S2 = P*P + Q2, S = sqrt(S2), T2 = PP1*PP1 + Q2, T = sqrt(T2);
  double SignP = (P > 0.0) ? 1.0 : -1.0;
  double Q, Q6, P4, PP14, T3, T4, T5, T6, T7, T8, S3, S4, S5, S6, S7, S8, LogFac, atanFac;
  switch (p) {
    case 0: *I0 = 1.0; *I1 = 0.5; return;
    case 1:
      S3 = S2*S; T3 = T2*T;
      if (Q2 < 1.0e-8) *I0 = SignP*(T2 - S2)/2.0;
      else { LogFac = log((PP1 + T)/(P + S)); *I0 = (T + P*(T - S) + Q2*LogFac)/2.0; }
      *I1 = -P*(*I0) + (T3 - S3)/3.0; return;
    case -1:
      if (Q2 < 1.0e-8) *I0 = SignP*log(1.0 + 1.0/P);
      else *I0 = log((PP1 + T)/(P + S));
      *I1 = -P*(*I0) + (T - S); return;
    case 2:
      S4 = S2*S2; T4 = T2*T2;
      *I0 = (T2 + S2)/2.0 - 1.0/6.0;
      *I1 = -P*(*I0) + (T4 - S4)/4.0; return;
    case -2:
      if (Q2 < 1.0e-8) *I0 = 1.0/(P*PP1);
      else { Q = sqrt(Q2); *I0 = (atan(PP1/Q) - atan(P/Q))/Q; }
      *I1 = -P*(*I0) + log(T/S); return;
    case 3:
      S4 = S2*S2; S5 = S4*S; T4 = T2*T2; T5 = T4*T;
      if (Q2 < 1.0e-8) *I0 = SignP*(T4 - S4)/4.0;
      else { LogFac = log((PP1 + T)/(P + S));
             *I0 = (PP1*(2.0*PP12 + 5.0*Q2)*T - P*(2.0*P2 + 5.0*Q2)*S + 3.0*Q4*LogFac)/8.0; }
      *I1 = -P*(*I0) + (T5 - S5)/5.0; return;
    case -3:
      if (Q2 < 1.0e-8) *I0 = SignP*(1.0/(2.0*S2) - 1.0/(2.0*T2));
      else *I0 = (PP1/T - P/S)/Q2;
      *I1 = -P*(*I0) + 1.0/S - 1.0/T; return;
    case 4:
      T4 = T2*T2; T6 = T4*T2; S4 = S2*S2; S6 = S4*S2;
      *I0 = (T4 + S4 + T2*S2)/3.0 - (T2 + S2)/6.0 + 1.0/30.0;
      *I1 = -P*(*I0) + (T6 - S6)/6.0; return;
    case -4:
      if (Q2 < 1.0e-6) { S3 = S*S2; T3 = T*T2; *I0 = SignP*(1.0/(3.0*S3) - 1.0/(3.0*T3)); }
      else { Q = sqrt(Q2); atanFac = (atan(PP1/Q) - atan(P/Q))/Q;
             *I0 = ((Q2 - P*PP1)/(S2*T2) + atanFac)/(2.0*Q2); }
      *I1 = -P*(*I0) + 1.0/(2.0*S2) - 1.0/(2.0*T2); return;
    case 5:
      P4 = P2*P2; PP14 = PP12*PP12; Q6 = Q4*Q2;
      S6 = S2*S2*S2; S7 = S6*S; T6 = T2*T2*T2; T7 = T6*T;
      if (Q2 < 1.0e-8) *I0 = SignP*(T6 - S6)/6.0;
      else { LogFac = log((PP1 + T)/(P + S));
             *I0 = (PP1*(8.0*PP14 + 26.0*PP12*Q2 + 33.0*Q4)*T
                    - P*(8.0*P4 + 26.0*P2*Q2 + 33.0*Q4)*S + 15.0*Q6*LogFac)/48.0; }
      *I1 = -P*(*I0) + (T7 - S7)/7.0; return;
    case -5:
      if (Q2 < 1.0e-4) { S3 = S2*S; S4 = S3*S; T3 = T2*T; T4 = T3*T;
                         *I0 = SignP*(1.0/(4.0*S4) - 1.0/(4.0*T4)); }
      else { S3 = S*S2; T3 = T*T2;
             *I0 = (PP1*(2.0*PP12 + 3.0*Q2)*S3 - P*(2.0*P2 + 3.0*Q2)*T3)/(3.0*Q4*S3*T3); }
      *I1 = -P*(*I0) + 1.0/(3.0*S3) - 1.0/(3.0*T3); return;
    case 6:
      T4 = T2*T2; T6 = T4*T2; T8 = T6*T2; S4 = S2*S2; S6 = S4*S2; S8 = S6*S2;
      *I0 = (T6 + S6 + S2*T4 + T2*S4)/4.0 - (3.0*T4 + 3.0*S4 + 4.0*S2*T2)/20.0
            + (T2 + S2)/20.0 - 1.0/140.0;
      *I1 = -P*(*I0) + (T8 - S8)/8.0; return;
    case -6:
      if (Q2 < 1.0e-4) { S4 = S2*S2; S5 = S4*S; T4 = T2*T2; T5 = T4*T;
                         *I0 = SignP*(1.0/(5.0*S5) - 1.0/(5.0*T5)); }
      else { Q = sqrt(Q2); Q4 = Q2*Q2; T4 = T2*T2; S4 = S2*S2;
             atanFac = (atan(PP1/Q) - atan(P/Q))/Q;
             *I0 = (3.0*PP1/T2 + 2.0*Q2*PP1/T4 - 3.0*P/S2 - 2.0*Q2*P/S4 + 3*atanFac)/(8.0*Q4); }
      *I1 = -P*(*I0) + 1.0/(4.0*S4) - 1.0/(4.0*T4); return;
    default:
      fprintf(stderr, "oracle: qfp_integral p=%d\n", p); abort();
  }
}

static const double fact_table[7] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 620.0};

static void get_script_jl(int WhichK, zdouble KParam, double Alpha, double Beta, double Gamma2,
                          int nMin, int nMax, zdouble *J, zdouble *L)
{
  double I0, I1;
  if (WhichK == OTD_RP) {
    double p = creal(KParam), AlphaP = pow(Alpha, p);
    if (p == 0) { I0 = 1.0; I1 = 0.5; } else qfp_integral(Beta, Gamma2, (int)p, &I0, &I1);
    for (int n = nMin; n <= nMax; n++) { J[n] = AlphaP*I0/(1.0 + n + p); L[n] = AlphaP*I1/(1.0 + n + p); }
  } else if (WhichK == OTD_HIGHK_HELMHOLTZ) {
    for (int n = nMin; n <= nMax; n++) {
      zdouble f = cpow(-I*KParam*Alpha, -(double)n);
      qfp_integral(Beta, Gamma2, -(n+1), &I0, &I1);
      J[n] = fact_table[n-1]*f*I0/Alpha; L[n] = fact_table[n-1]*f*I1/Alpha;
    }
  } else if (WhichK == OTD_HIGHK_GRADHELMHOLTZ) {
    double A3 = Alpha*Alpha*Alpha;
    for (int n = nMin; n <= nMax; n++) {
      zdouble f = cpow(-I*KParam*Alpha, -(double)(n-2));
      qfp_integral(Beta, Gamma2, -(n+1), &I0, &I1);
      J[n] = -(n-1)*fact_table[n-3]*f*I0/A3; L[n] = -(n-1)*fact_table[n-3]*f*I1/A3;
    }
  }
}

/* relative exponential, TaylorDuffy.cc:840-893 */
static zdouble exprel_v3(int n, zdouble Z)
{
  int m;
  if (cimag(Z) == 0.0) {
    double X = creal(Z), Term, Sum;
    if (fabs(X) < 0.1) {
      Sum = 1.0;
      for (Term = 1.0, m = 1; m < 100; m++) {
        Term *= X/((double)(m+n));
        Sum += Term;
        if (fabs(Term) < 1.0e-8*fabs(Sum)) break;
      }
      return Sum;
    }
    Sum = exp(X);
    for (Term = 1.0, m = 0; m < n; m++) { Sum -= Term; Term *= X/((double)(m+1)); }
    return Sum/Term;
  } else {
    zdouble Term, Sum;
    if (cabs(Z) < 0.1) {
      for (Sum = Term = 1.0, m = 1; m < 100; m++) {
        Term *= Z/((double)(m+n));
        Sum += Term;
        double nt = creal(Term)*creal(Term) + cimag(Term)*cimag(Term);
        double ns = creal(Sum)*creal(Sum) + cimag(Sum)*cimag(Sum);
        if (nt < 1.0e-8*1.0e-8*ns) break;
      }
      return Sum;
    }
    Sum = cexp(Z);
    for (Term = 1.0, m = 0; m < n; m++) { Sum -= Term; Term *= Z/((double)(m+1)); }
    return Sum/Term;
  }
}

static void get_script_k(int WhichK, zdouble KParam, double X, int nMin, int nMax, zdouble *K)
{
  if (WhichK == OTD_RP) {
    double P = creal(KParam), XP = pow(X, P);
    for (int n = nMin; n <= nMax; n++) K[n] = XP/(1.0 + P + (double)n);
  } else if (WhichK == OTD_HELMHOLTZ) {
    zdouble IK = I*KParam, IKX = IK*X, e = cexp(IKX);
    for (int n = nMin; n <= nMax; n++) K[n] = e*exprel_v3(n, -IKX)/(n*X);
  } else if (WhichK == OTD_GRADHELMHOLTZ) {
    zdouble IK = I*KParam, IKX = IK*X, e = cexp(IKX), ER[10];
    if ((nMin-2) < 0 || (nMax-1) >= 10) {
      for (int n = 0; n < nMin; n++) K[n] = 0.0;
      nMin = 2;
    }
    for (int n = nMin-2; n <= nMax-1; n++) ER[n] = exprel_v3(n, -IKX);
    for (int n = nMin; n <= nMax; n++)
      K[n] = e*(IK*ER[n-1]/((double)n - 1.0) - ER[n-2]/(((double)n - 2.0)*X))/(X*X);
  }
}

/* ------------------------------------------------------------------ */
/* Upsilon coefficient tables (TaylorDuffy.cc:940-1262).  C = {C2010,  */
/* C2001, C0210, C0201}, M symmetric 4x4 (upper triangle used), V[4],  */
/* S.                                                                  */
/* ------------------------------------------------------------------ */
static void cmvs_to_upsilon(int WhichCase, const double *C, const double *M, const double *V,
                            double S, double U[NREG][NWP][NMON])
{
  memset(U, 0, sizeof(double)*NREG*NWP*NMON);
  double C2010 = C[0], C2001 = C[1], C0210 = C[2], C0201 = C[3];
  double M11 = M[0], M12 = M[4], M13 = M[8], M14 = M[12];
  double M22 = M[5], M23 = M[9], M24 = M[13];
  double M33 = M[10], M34 = M[14];
  double M44 = M[15];
  double V1 = V[0], V2 = V[1], V3 = V[2], V4 = V[3];

  if (WhichCase == OTD_COMMONTRIANGLE) {
    double u00 = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
                 + M34/4 + M44/12 + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3 + S;
    U[0][0][T1] = u00;
    U[0][1][TY1] = M12/3 + M14/3 + M22/6 + M23/3 + M24/3 + M34/3 + M44/6 + V2/2 + V4/2;
    U[0][1][T1]  = -(2*M11)/3 - (5*M12)/6 - (4*M13)/3 - (5*M14)/6 - M22/3 - (5*M23)/6
                   - (2*M24)/3 - (2*M33)/3 - (5*M34)/6 - M44/3 - (3*V1)/2 - V2 - (3*V3)/2
                   - V4 - 2*S;
    U[0][2][TY12] = M22/4 + M44/4;
    U[0][2][TY1]  = -M12/2 - M14 - M22/2 - M23 - M24 - M34/2 - M44/2 - V2 - V4;
    U[0][2][T1]   = (3*M11)/4 + M12 + M13 + M14 + M22/2 + M23 + M24
                    + (3*M33)/4 + M34 + M44/2 + V1 + V2 + V3 + V4 + S;
    U[0][3][TY12] = -M22/2 - M44/2;
    U[0][3][TY1]  = M14 + M22/2 + M23 + M24 + M44/2 + V2/2 + V4/2;
    U[0][3][T1]   = -M11/2 - M12/2 - M14/2 - M22/3 - M23/2 - (2*M24)/3
                    - M33/2 - M34/2 - M44/3 - V1/6 - V2/3 - V3/6 - V4/3;
    U[0][4][TY12] = M22/4 + M44/4;
    U[0][4][TY1]  = M12/6 - M14/3 - M22/6 - M23/3 - M24/3 + M34/6 - M44/6;
    U[0][4][T1]   = M11/6 + M12/12 - M13/6 + M14/12 + M22/12 + M23/12 + M24/6
                    + M33/6 + M34/12 + M44/12;

    U[1][0][T1] = u00;
    U[1][1][TY1] = -M11/3 - M12/2 - (2*M13)/3 - M14/2 - M22/6 - M23/2 - M24/3 - M33/3
                   - M34/2 - M44/6 - V1/2 - V2/2 - V3/2 - V4/2;
    U[1][1][T1]  = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3 - M33/3
                   - M34/3 - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
    U[1][2][TY12] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
    U[1][2][TY1]  = M11/2 + M12/2 + M13 + M14 + M23 + M24 + M33/2 + M34/2
                    + V1 + V2 + V3 + V4;
    U[1][2][T1]   = M22/4 + M44/4 + S;
    U[1][3][TY12] = -M11/2 - M12 - M22/2 - M33/2 - M34 - M44/2;
    U[1][3][TY1]  = M12/2 - M14/2 + M22/2 - M23/2 - M24 + M34/2 + M44/2
                    - V1/2 - V2/2 - V3/2 - V4/2;
    U[1][3][T1]   = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
    U[1][4][TY12] = M11/4 + M12/2 + M22/4 + M33/4 + M34/2 + M44/4;
    U[1][4][TY1]  = -M11/6 - M12/2 - M13/3 - M22/3 + M24/3 - M33/6 - M34/2 - M44/3;
    U[1][4][T1]   = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6
                    + M33/12 + M34/12 + M44/6;

    U[2][0][T1] = M11/4 + M12/4 + M13/2 + M14/4 + M22/12 + M23/4 + M24/6 + M33/4
                  + M34/4 + M44/12 + S + (2*V1)/3 + V2/3 + (2*V3)/3 + V4/3;
    U[2][1][TY1] = -M11/3 - M12/6 - (2*M13)/3 - M14/6 - M23/6 - M33/3 - M34/6 - V1/2 - V3/2;
    U[2][1][T1]  = -M11/3 - M12/3 - (2*M13)/3 - M14/3 - M22/6 - M23/3 - M24/3
                   - M33/3 - M34/3 - M44/6 - 2*S - V1 - V2/2 - V3 - V4/2;
    U[2][2][TY12] = M11/4 + M33/4;
    U[2][2][TY1]  = M11/2 + M12/2 + M13 + M33/2 + M34/2 + V1 + V3;
    U[2][2][T1]   = M22/4 + M44/4 + S;
    U[2][3][TY12] = -M11/2 - M33/2;
    U[2][3][TY1]  = -M12/2 + M14/2 + M23/2 - M34/2 - V1/2 - V3/2;
    U[2][3][T1]   = -M22/3 + M24/3 - M44/3 + V1/3 + V2/6 + V3/3 + V4/6;
    U[2][4][TY12] = M11/4 + M33/4;
    U[2][4][TY1]  = -M11/6 + M12/6 - M13/3 - M14/3 - M23/3 - M33/6 + M34/6;
    U[2][4][T1]   = M11/12 + M12/12 + M13/6 + M14/12 + M22/6 + M23/12 - M24/6
                    + M33/12 + M34/12 + M44/6;
  } else if (WhichCase == OTD_COMMONEDGE) {
    double u00 = C2010/4 + M11/6 + M13/3 + M33/6 + S + V1/2 + V3/2;
    double u10 = C2001/3 + M12/2 + M14/2 + M23/2 + M34/2 - S + V2 + V4;
    double u20 = C0210/2 + M22/2 + M24 + M44/2 - V1/2 - V2 - V3/2 - V4;
    double u30 = C0201 - M11/6 - M12/2 - M13/3 - M14/2 - M22/2 - M23/2 - M24 - M33/6 - M34/2 - M44/2;
    double u40 = -C0201 - C0210/2 - C2001/3 - C2010/4;

    /* d = 0 */
    U[0][0][T1] = u00;
    U[0][1][TY1Y2] = M12/2 + M23/2 + V2;
    U[0][1][TY1]   = -C2001/3 - C2010/3 - M12/2 - M13/2 - M14/2 - M23/2 - M33/2 - M34/2 - V2 - V3 - V4;
    U[0][1][T1]    = u10;
    U[0][2][TY12Y22] = C0210/2 + M22/2;
    U[0][2][TY12Y2]  = -C0210 - M22 - M23 - M24;
    U[0][2][TY12]    = C0210/2 + M22/2 + M23 + M24 + M33/2 + M34 + M44/2;
    U[0][2][TY1Y2]   = C0210 + M22 + M24 - V2;
    U[0][2][TY1]     = -C0210 - M22 - M23 - 2*M24 - M34 - M44 + V2 + V3 + V4;
    U[0][2][T1]      = u20;
    U[0][3][TY13Y22] = -C0201 - C0210;
    U[0][3][TY13Y2]  = 2*C0201 + 2*C0210;
    U[0][3][TY13]    = -C0201 - C0210;
    U[0][3][TY12Y22] = C0201 - M22/2;
    U[0][3][TY12Y2]  = -4*C0201 - 2*C0210 + M22 + M23 + M24;
    U[0][3][TY12]    = 3*C0201 + 2*C0210 - M22/2 - M23 - M24 - M33/2 - M34 - M44/2;
    U[0][3][TY1Y2]   = 2*C0201 - M12/2 - M22 - M23/2 - M24;
    U[0][3][TY1]     = -3*C0201 - C0210 + M12/2 + M13/2 + M14/2 + M22 + (3*M23)/2 + 2*M24
                       + M33/2 + (3*M34)/2 + M44;
    U[0][3][T1]      = u30;
    U[0][4][TY13Y22] = C0201 + C0210;
    U[0][4][TY13Y2]  = -2*C0201 - 2*C0210;
    U[0][4][TY13]    = C0201 + C0210;
    U[0][4][TY12Y22] = -C0201 - C0210/2;
    U[0][4][TY12Y2]  = 4*C0201 + 3*C0210;
    U[0][4][TY12]    = -3*C0201 - (5*C0210)/2;
    U[0][4][TY1Y2]   = -2*C0201 - C0210;
    U[0][4][TY1]     = 3*C0201 + 2*C0210 + C2001/3 + C2010/3;
    U[0][4][T1]      = u40;

    /* d = 1 */
    U[1][0][T1] = u00;
    U[1][1][TY1Y2] = C2001/3 + M14/2 + M34/2 + V4;
    U[1][1][TY1]   = -C2001/3 - (2*C2010)/3 - M11/2 - M12/2 - M13/2 - M14/2 - M23/2 - M34/2 - V1 - V2 - V4;
    U[1][1][T1]    = u10;
    U[1][2][TY12Y22] = M44/2;
    U[1][2][TY12Y2]  = -C2001 - M14 - M24 - M44;
    U[1][2][TY12]    = C0210/2 + C2001 + C2010/2 + M11/2 + M12 + M14 + M22/2 + M24 + M44/2;
    U[1][2][TY1Y2]   = M24 + M44 - V4;
    U[1][2][TY1]     = -C0210 - C2001 - M12 - M14 - M22 - 2*M24 - M44 + V1 + V2 + V4;
    U[1][2][T1]      = u20;
    U[1][3][TY13Y2]  = C0201 + C2001;
    U[1][3][TY13]    = -C0201 - C2001;
    U[1][3][TY12Y22] = -M44/2;
    U[1][3][TY12Y2]  = -2*C0201 + M14 + M24 + M44;
    U[1][3][TY12]    = 3*C0201 + C2001 - M11/2 - M12 - M14 - M22/2 - M24 - M44/2;
    U[1][3][TY1Y2]   = C0201 - M14/2 - M24 - M34/2 - M44;
    U[1][3][TY1]     = -3*C0201 + M11/2 + (3*M12)/2 + M13/2 + (3*M14)/2 + M22 + M23/2 + 2*M24
                       + M34/2 + M44;
    U[1][3][T1]      = u30;
    U[1][4][TY13Y2]  = -C0201 - C2001;
    U[1][4][TY13]    = C0201 + C2001;
    U[1][4][TY12Y2]  = 2*C0201 + C2001;
    U[1][4][TY12]    = -3*C0201 - C0210/2 - 2*C2001 - C2010/2;
    U[1][4][TY1Y2]   = -C0201 - C2001/3;
    U[1][4][TY1]     = 3*C0201 + C0210 + (4*C2001)/3 + (2*C2010)/3;
    U[1][4][T1]      = u40;

    /* d = 2 */
    U[2][0][T1] = u00;
    U[2][1][TY1Y2] = -C2001/3 - C2010/3 - M13/2 - M14/2 - M33/2 - M34/2 - V3 - V4;
    U[2][1][TY1]   = -M12/2 - M23/2 - V2;
    U[2][1][T1]    = u10;
    U[2][2][TY12Y22] = M33/2 + M34 + M44/2;
    U[2][2][TY12Y2]  = M23 + M24;
    U[2][2][TY12]    = C0210/2 + M22/2;
    U[2][2][TY1Y2]   = -M23 - M24 - M34 - M44 + V3 + V4;
    U[2][2][TY1]     = -C0210 - M22 - M24 + V2;
    U[2][2][T1]      = u20;
    U[2][3][TY13Y2]  = -C0201 - C0210;
    U[2][3][TY12Y22] = -M33/2 - M34 - M44/2;
    U[2][3][TY12Y2]  = 2*C0201 + 2*C0210 - M23 - M24;
    U[2][3][TY12]    = C0201 - M22/2;
    U[2][3][TY1Y2]   = -C0201 - C0210 + M13/2 + M14/2 + M23 + M24 + M33/2 + (3*M34)/2 + M44;
    U[2][3][TY1]     = -2*C0201 + M12/2 + M22 + M23/2 + M24;
    U[2][3][T1]      = u30;
    U[2][4][TY13Y2]  = C0201 + C0210;
    U[2][4][TY12Y2]  = -2*C0201 - 2*C0210;
    U[2][4][TY12]    = -C0201 - C0210/2;
    U[2][4][TY1Y2]   = C0201 + C0210 + C2001/3 + C2010/3;
    U[2][4][TY1]     = 2*C0201 + C0210;
    U[2][4][T1]      = u40;

    /* d = 3 */
    U[3][0][T1] = u00;
    U[3][1][TY1Y2] = (-2*C2010)/3 - M11/2 - M12/2 - M13/2 - M23/2 - V1 - V2;
    U[3][1][TY1]   = -C2001/3 - M14/2 - M34/2 - V4;
    U[3][1][T1]    = u10;
    U[3][2][TY12Y22] = C0210/2 + C2010/2 + M11/2 + M12 + M22/2;
    U[3][2][TY12Y2]  = C2001 + M14 + M24;
    U[3][2][TY12]    = M44/2;
    U[3][2][TY1Y2]   = -C0210 - C2001 - M12 - M14 - M22 - M24 + V1 + V2;
    U[3][2][TY1]     = -M24 - M44 + V4;
    U[3][2][T1]      = u20;
    U[3][3][TY13Y22] = -C0201 - C2001;
    U[3][3][TY12Y22] = C0201 + C2001 - M11/2 - M12 - M22/2;
    U[3][3][TY12Y2]  = 2*C0201 - M14 - M24;
    U[3][3][TY12]    = -M44/2;
    U[3][3][TY1Y2]   = -2*C0201 + M11/2 + (3*M12)/2 + M13/2 + M14 + M22 + M23/2 + M24;
    U[3][3][TY1]     = -C0201 + M14/2 + M24 + M34/2 + M44;
    U[3][3][T1]      = u30;
    U[3][4][TY13Y22] = C0201 + C2001;
    U[3][4][TY12Y22] = -C0201 - C0210/2 - C2001 - C2010/2;
    U[3][4][TY12Y2]  = -2*C0201 - C2001;
    U[3][4][TY1Y2]   = 2*C0201 + C0210 + C2001 + (2*C2010)/3;
    U[3][4][TY1]     = C0201 + C2001/3;
    U[3][4][T1]      = u40;

    /* d = 4 */
    U[4][0][T1] = u00;
    U[4][1][TY1Y2] = -C2010/3 - M13/2 - M33/2 - V3;
    U[4][1][TY1]   = -C2001/3 - M14/2 - M34/2 - V4;
    U[4][1][T1]    = u10;
    U[4][2][TY12Y22] = M33/2;
    U[4][2][TY12Y2]  = M34;
    U[4][2][TY12]    = M44/2;
    U[4][2][TY1Y2]   = -M23 - M34 + V3;
    U[4][2][TY1]     = -M24 - M44 + V4;
    U[4][2][T1]      = u20;
    U[4][3][TY12Y22] = -M33/2;
    U[4][3][TY12Y2]  = -M34;
    U[4][3][TY12]    = -M44/2;
    U[4][3][TY1Y2]   = -C0210 + M13/2 + M23 + M33/2 + M34;
    U[4][3][TY1]     = -C0201 + M14/2 + M24 + M34/2 + M44;
    U[4][3][T1]      = u30;
    U[4][4][TY1Y2]   = C0210 + C2010/3;
    U[4][4][TY1]     = C0201 + C2001/3;
    U[4][4][T1]      = u40;

    /* d = 5 */
    U[5][0][T1] = u00;
    U[5][1][TY1Y2] = (-2*C2010)/3 - M11/2 - M13/2 - V1;
    U[5][1][TY1]   = -M12/2 - M23/2 - V2;
    U[5][1][T1]    = u10;
    U[5][2][TY12Y22] = C2010/2 + M11/2;
    U[5][2][TY12Y2]  = M12;
    U[5][2][TY12]    = C0210/2 + M22/2;
    U[5][2][TY1Y2]   = -C2001 - M12 - M14 + V1;
    U[5][2][TY1]     = -C0210 - M22 - M24 + V2;
    U[5][2][T1]      = u20;
    U[5][3][TY12Y22] = C2001 - M11/2;
    U[5][3][TY12Y2]  = -M12;
    U[5][3][TY12]    = C0201 - M22/2;
    U[5][3][TY1Y2]   = M11/2 + M12 + M13/2 + M14;
    U[5][3][TY1]     = -2*C0201 + M12/2 + M22 + M23/2 + M24;
    U[5][3][T1]      = u30;
    U[5][4][TY12Y22] = -C2001 - C2010/2;
    U[5][4][TY12]    = -C0201 - C0210/2;
    U[5][4][TY1Y2]   = C2001 + (2*C2010)/3;
    U[5][4][TY1]     = 2*C0201 + C0210;
    U[5][4][T1]      = u40;
  } else {
    /* common vertex */
    U[0][0][T1] = S;
    U[0][1][T1] = V1; U[0][1][TY1] = V2; U[0][1][TY2] = V3; U[0][1][TY2Y3] = V4;
    U[0][2][T1] = M11/2.0; U[0][2][TY1] = M12; U[0][2][TY12] = M22/2.0;
    U[0][2][TY2] = M13; U[0][2][TY1Y2] = M23; U[0][2][TY22] = M33/2.0;
    U[0][2][TY2Y3] = M14; U[0][2][TY1Y2Y3] = M24; U[0][2][TY22Y3] = M34; U[0][2][TY22Y32] = M44/2.0;
    U[0][3][TY2] = C2010; U[0][3][TY12Y2] = C0210; U[0][3][TY2Y3] = C2001; U[0][3][TY12Y2Y3] = C0201;

    U[1][0][T1] = S;
    U[1][1][T1] = V3; U[1][1][TY1] = V4; U[1][1][TY2] = V1; U[1][1][TY2Y3] = V2;
    U[1][2][T1] = M33/2.0; U[1][2][TY1] = M34; U[1][2][TY12] = M44/2.0;
    U[1][2][TY2] = M13; U[1][2][TY1Y2] = M14; U[1][2][TY22] = M11/2.0;
    U[1][2][TY2Y3] = M23; U[1][2][TY1Y2Y3] = M24; U[1][2][TY22Y3] = M12; U[1][2][TY22Y32] = M22/2.0;
    U[1][3][TY22] = C2010; U[1][3][TY1Y22] = C2001; U[1][3][TY22Y32] = C0210; U[1][3][TY1Y22Y32] = C0201;
  }
}

/* ComputeGeometricParameters (TaylorDuffy.cc:1265-1528) */
static void compute_geometry(const OTDArgs *Args, tdw *W)
{
  double *V1 = Args->V1, *V2 = Args->V2, *V3 = Args->V3, *V2P = Args->V2P, *V3P = Args->V3P;
  double *Q = Args->Q, *QP = Args->QP;
  if (Args->WhichCase == OTD_COMMONTRIANGLE) { V2P = V2; V3P = V3; }
  else if (Args->WhichCase == OTD_COMMONEDGE) V2P = V2;
  double A[3], B[3], AP[3], BP[3], L[3];
  sub3(V2, V1, A); sub3(V3, V2, B); sub3(V2P, V1, AP); sub3(V3P, V2P, BP);
  W->A2 = dot3(A, A); W->AdB = dot3(A, B); W->B2 = dot3(B, B);
  if (Args->WhichCase == OTD_COMMONEDGE) {
    sub3(BP, B, L);
    W->BP2 = dot3(BP, BP); W->AdBP = dot3(A, BP); W->L2 = dot3(L, L);
    W->AdL = dot3(A, L); W->BPdL = dot3(BP, L);
  } else if (Args->WhichCase == OTD_COMMONVERTEX) {
    W->AP2 = dot3(AP, AP); W->BP2 = dot3(BP, BP); W->AdAP = dot3(A, AP); W->AdBP = dot3(A, BP);
    W->BdAP = dot3(B, AP); W->BdBP = dot3(B, BP); W->APdBP = dot3(AP, BP);
  }
  int NumRegions = (Args->WhichCase == OTD_COMMONVERTEX) ? 2 : (Args->WhichCase == OTD_COMMONEDGE ? 6 : 3);
  double D[3] = {0,0,0}, DP[3] = {0,0,0}, DeltaD[3] = {0,0,0}, Scr[3], Scr2[3];
  if (Q && QP) { sub3(V1, Q, D); sub3(V1, QP, DP); sub3(QP, Q, DeltaD); }
  for (int np = 0; np < NP; np++) {
    if (!W->NeedP[np]) continue;
    double C[4] = {0,0,0,0}, M[16], V[4] = {0,0,0,0}, S = 0.0;
    memset(M, 0, sizeof(M));
    if (np == OTD_UNITY) S = 1;
    else if (np == OTD_PMCHWG1) {
      M[0 + 2*4] = dot3(A, AP); M[0 + 3*4] = dot3(A, BP);
      M[1 + 2*4] = dot3(B, AP); M[1 + 3*4] = dot3(B, BP);
      V[0] = dot3(A, DP); V[1] = dot3(B, DP); V[2] = dot3(D, AP); V[3] = dot3(D, BP);
      S = dot3(D, DP);
    } else if (np == OTD_PMCHWC) {
      M[0 + 2*4] = -dot3(cross3(A, AP, Scr), DeltaD);
      M[0 + 3*4] = -dot3(cross3(A, BP, Scr), DeltaD);
      M[1 + 2*4] = -dot3(cross3(B, AP, Scr), DeltaD);
      M[1 + 3*4] = -dot3(cross3(B, BP, Scr), DeltaD);
      cross3(DP, D, Scr2);
      V[0] = -dot3(A, Scr2); V[1] = -dot3(B, Scr2); V[2] = dot3(AP, Scr2); V[3] = dot3(BP, Scr2);
    }
    cmvs_to_upsilon(Args->WhichCase, C, M, V, S, W->U[np]);
    int nMin = NWP - 1, nMax = -1;
    for (int n = 0; n < NWP; n++)
      for (int nr = 0; nr < NumRegions; nr++)
        for (int nm = 0; nm < NMON; nm++)
          if (fabs(W->U[np][nr][n][nm]) > 1.0e-8) { if (n < nMin) nMin = n; if (n > nMax) nMax = n; }
    W->nMin[np] = nMin; W->nMax[np] = nMax;
  }
}

/* TaylorDuffySum (TaylorDuffy.cc:152-258) */
static void td_sum(unsigned ndim, const double *y, void *parms, unsigned nfun, double *f)
{
  (void)ndim;
  tdw *W = (tdw *)parms;
  W->nCalls++;
  int NumRegions, nOffset;
  double yIn, Jac;
  int TI = W->TwiceIntegrable;
  if (W->WhichCase == OTD_COMMONTRIANGLE) { NumRegions = 3; nOffset = 1; yIn = TI ? 0.0 : y[0]; Jac = 1.0; }
  else if (W->WhichCase == OTD_COMMONEDGE) { NumRegions = 6; nOffset = 2; yIn = TI ? 0.0 : y[1]; Jac = y[0]; }
  else { NumRegions = 2; nOffset = 3; yIn = TI ? 0.0 : y[2]; Jac = y[1]; }
  if (Jac == 0.0) { memset(f, 0, nfun*sizeof(double)); return; }
  double X[NREG], A[NREG], B[NREG], G2[NREG];
  if (TI) get_abg2(W, y, A, B, G2); else get_x(W, y, X);
  double P[NP][NREG][NWP][3];
  for (int np = 0; np < NP; np++) if (W->NeedP[np]) get_script_p(W, np, y, P[np]);
  zdouble J[NREG][8], L[NREG][8], K[NREG][8];
  for (int npk = 0; npk < W->NumPKs; npk++) {
    int np = W->PIndex[npk], nMin = W->nMin[np], nMax = W->nMax[np];
    if (TI) for (int d = 0; d < NumRegions; d++)
      get_script_jl(W->KIndex[npk], W->KParam[npk], A[d], B[d], G2[d], nMin+nOffset, nMax+nOffset, J[d], L[d]);
    else for (int d = 0; d < NumRegions; d++)
      get_script_k(W->KIndex[npk], W->KParam[npk], X[d], nMin+nOffset, nMax+nOffset, K[d]);
    zdouble s = 0.0;
    if (TI) {
      for (int n = nMin; n <= nMax; n++)
        for (int d = 0; d < NumRegions; d++)
          s += P[np][d][n][0]*J[d][n+nOffset] + P[np][d][n][1]*L[d][n+nOffset];
    } else {
      for (int n = nMin; n <= nMax; n++)
        for (int d = 0; d < NumRegions; d++)
          s += (P[np][d][n][0] + yIn*P[np][d][n][1])*K[d][n+nOffset];
    }
    s *= Jac/(4.0*M_PI);
    f[2*npk] = creal(s); f[2*npk+1] = cimag(s);
  }
}

void o_taylor_duffy_init(OTDArgs *A)
{
  memset(A, 0, sizeof(*A));
  A->AbsTol = 0.0;
  A->RelTol = 1.0e-10;
  A->MaxEval = 1000;
}

/* TaylorDuffy driver (TaylorDuffy.cc:277-367) */
void o_taylor_duffy(OTDArgs *Args)
{
  /* environment overrides (TaylorDuffy.cc:294-296) */
  const char *ev;
  if ((ev = getenv("SCUFF_TAYLORDUFFY_ABSTOL"))) Args->AbsTol = atof(ev);
  if ((ev = getenv("SCUFF_TAYLORDUFFY_RELTOL"))) Args->RelTol = atof(ev);
  if ((ev = getenv("SCUFF_TAYLORDUFFY_MAXEVAL"))) Args->MaxEval = atoi(ev);
  tdw *W = (tdw *)calloc(1, sizeof(tdw));
  W->WhichCase = Args->WhichCase; W->NumPKs = Args->NumPKs;
  W->PIndex = Args->PIndex; W->KIndex = Args->KIndex; W->KParam = Args->KParam;
  for (int npk = 0; npk < Args->NumPKs; npk++) W->NeedP[Args->PIndex[npk]] = 1;
  compute_geometry(Args, W);
  int TI = 1;
  for (int npk = 0; TI && npk < Args->NumPKs; npk++)
    if (Args->KIndex[npk] == OTD_HELMHOLTZ || Args->KIndex[npk] == OTD_GRADHELMHOLTZ) TI = 0;
  W->TwiceIntegrable = TI;
  const double Lower[3] = {0, 0, 0}, Upper[3] = {1, 1, 1};
  int dim = 4 - Args->WhichCase - TI;
  double *R = (double *)Args->Result, *E = (double *)Args->Error;
  if (dim == 0) td_sum(0, NULL, W, 2*Args->NumPKs, R);
  else o_pcubature(2*Args->NumPKs, td_sum, W, dim, Lower, Upper, Args->MaxEval,
                   Args->AbsTol, Args->RelTol, R, E);
  Args->nCalls = W->nCalls;
  free(W);
}
