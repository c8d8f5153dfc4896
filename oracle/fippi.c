/*
 * fippi.c -- oracle: frequency-independent panel-panel integrals
 * (FIPPIs) and their cache.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  ncv==0  3-D adaptive cubature of the 33 moments with the innermost
 *          v' integral done in closed form
 *          (libs/libscuff/QIFIPPICubature.cc:187-452)
 *  ncv>0   Taylor-Duffy "V2P0" 1-D (common edge) or 2-D (common
 *          vertex) adaptive cubature (QIFIPPITaylorDuffyV2P0.cc:91-312,
 *          :329-478 and the polynomial tables at :484-900)
 *  cache   key = 15 floats of canonically ordered vertex differences,
 *          exact-bit match (FIPPICache.cc:120-187)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static void sub3(const double *a, const double *b, double *c) { c[0]=a[0]-b[0]; c[1]=a[1]-b[1]; c[2]=a[2]-b[2]; }
static double *cross3(const double *a, const double *b, double *c)
{ c[0]=a[1]*b[2]-a[2]*b[1]; c[1]=a[2]*b[0]-a[0]*b[2]; c[2]=a[0]*b[1]-a[1]*b[0]; return c; }

/* ------------------------------------------------------------------ */
/* ncv = 0: 3-D cubature integrand (QIFIPPICubature.cc:187-325)        */
/* x = V0 + u A + v B (v = u*x1), x' = V0' + u' A' + v' B', the v'      */
/* integral over [0,u'] is done analytically.                          */
/* ------------------------------------------------------------------ */
typedef struct { const double *V0; double A[3], B[3]; const double *V0P; double AP[3], BP[3]; size_t nCalls; } cfd;

static void cfd_integrand(unsigned ndim, const double *x, void *params, unsigned fdim, double *fval)
{
  (void)ndim; (void)fdim;
  cfd *D = (cfd *)params;
  D->nCalls++;
  const double *V0 = D->V0, *A = D->A, *B = D->B, *V0P = D->V0P, *AP = D->AP, *BP = D->BP;
  double u = x[0], v = u*x[1], up = x[2];
  double X[3], Y[3], Z[3], BPdY = 0, Y2 = 0, a2 = 0;
  for (int m = 0; m < 3; m++) {
    X[m] = V0[m] + u*A[m] + v*B[m];
    Z[m] = V0P[m] + up*AP[m];
    Y[m] = X[m] - Z[m];
    a2 += BP[m]*BP[m];
    BPdY += BP[m]*Y[m];
    Y2 += Y[m]*Y[m];
  }
  double a = sqrt(a2), a3 = a*a2;
  double vp0 = -BPdY/a2, vp02 = vp0*vp0, vp03 = vp02*vp0, vp04 = vp03*vp0;
  double b2 = Y2/a2 - vp02;
  double OneRM3, vpRM3, OneRM1, vpRM1, OneR1, vpR1, OneR2, vpR2;
  if (b2 > 1.0e-10) {
    double S1 = sqrt(b2 + vp02);
    double S2 = sqrt(b2 + (vp0 + up)*(vp0 + up));
    double LogFac = log((S2 + (up + vp0))/(S1 + vp0));
    double Sum = vp0 + up, Sum3 = Sum*Sum*Sum, Sum4 = Sum3*Sum;
    OneRM3 = u*((up + vp0)/S2 - vp0/S1)/(a3*b2);
    vpRM3  = -vp0*OneRM3 + u*(1.0/S1 - 1.0/S2)/a3;
    OneRM1 = u*LogFac/a;
    vpRM1  = -vp0*OneRM1 + u*(S2 - S1)/a;
    OneR1  = u*a*0.5*(b2*LogFac + (up + vp0)*S2 - vp0*S1);
    vpR1   = -vp0*OneR1 + u*a*(S2*S2*S2 - S1*S1*S1)/3.0;
    OneR2  = u*a2*((Sum3 - vp03)/3.0 + up*b2);
    vpR2   = -vp0*OneR2 + u*a2*((Sum4 - vp04)/4.0 + up*(vp0 + 0.5*up)*b2);
  } else {
    double s = vp0 + up, s2 = s*s, s3 = s2*s, s4 = s3*s;
    OneRM3 = u*fabs(1.0/vp02 - 1.0/s2)/(2.0*a3);
    vpRM3  = -vp0*OneRM3 + u*fabs(1.0/vp0 - 1.0/s)/a3;
    OneRM1 = u*fabs(log(s/vp0))/a;
    vpRM1  = -vp0*OneRM1 + u*up/a;
    OneR1  = u*a*fabs(s2 - vp02)/2.0;
    OneR2  = u*a2*fabs(s3 - vp03)/3.0;
    vpR1   = -vp0*OneR1 + OneR2/a;
    vpR2   = -vp0*OneR2 + u*a2*fabs(s4 - vp04)/4.0;
  }
  double CPCT[3], CPLT[3];
  cross3(X, Z, CPCT);
  cross3(X, BP, CPLT);
  int nf = 0;
  for (int m = 0; m < 3; m++) fval[nf++] = Y[m]*OneRM3 - BP[m]*vpRM3;
  for (int m = 0; m < 3; m++) fval[nf++] = CPCT[m]*OneRM3 + CPLT[m]*vpRM3;
  const double I1[3] = {OneRM1, OneR1, OneR2}, Iv[3] = {vpRM1, vpR1, vpR2};
  for (int q = 0; q < 3; q++) {
    fval[nf++] = I1[q];
    fval[nf++] = up*I1[q];
    fval[nf++] = Iv[q];
    fval[nf++] = u*I1[q];
    fval[nf++] = u*up*I1[q];
    fval[nf++] = u*Iv[q];
    fval[nf++] = v*I1[q];
    fval[nf++] = v*up*I1[q];
    fval[nf++] = v*Iv[q];
  }
}

static void qifippi_cubature(double **Va, double **Vb, OQIFIPPIData *Q, size_t *nevals)
{
  cfd D;
  D.V0 = Va[0]; sub3(Va[1], Va[0], D.A); sub3(Va[2], Va[1], D.B);
  D.V0P = Vb[0]; sub3(Vb[1], Vb[0], D.AP); sub3(Vb[2], Vb[1], D.BP);
  D.nCalls = 0;
  double F[33], E[33];
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  o_hcubature(33, cfd_integrand, &D, 3, lo, hi, 1000000, 1.0e-10, 1.0e-6, F, E, NULL);
  if (nevals) *nevals = D.nCalls;
  memcpy(Q->xMxpRM3, F+0, 3*sizeof(double));
  memcpy(Q->xXxpRM3, F+3, 3*sizeof(double));
  memcpy(Q->uvupvpRM1, F+6, 9*sizeof(double));
  memcpy(Q->uvupvpR1, F+15, 9*sizeof(double));
  memcpy(Q->uvupvpR2, F+24, 9*sizeof(double));
}

/* ------------------------------------------------------------------ */
/* ncv = 1, 2: Taylor-Duffy V2P0 (QIFIPPITaylorDuffyV2P0.cc)           */
/* ------------------------------------------------------------------ */
typedef struct {
  double A[3], B[3], AP[3], BP[3];
  double V0xAP[3], V0xBP[3], AxV0[3], BxV0[3], AxAP[3], AxBP[3], BxAP[3], BxBP[3];
  double A2, B2, AP2, BP2, L2, AdB, AdAP, AdBP, AdL, BdAP, BdBP, APdBP, BPdL;
  double PCross[6][6][4][2], PXXEE[9][6][4][2];
  int Case;
  size_t nCalls;
} tdf;

static void pqr_abg2(double P, double Q, double R, double *A, double *B, double *G2)
{
  *A = sqrt(P);
  *B = Q/P;
  double RoP = R/P;
  *G2 = RoP - (*B)*(*B);
  if (fabs(*G2) < 1.0e-6*RoP) { *B = 0.0; *G2 = RoP; }
}

/* QIFIPPITaylorDuffyV2P0.cc:345-428 */
static void tdf_abg(const tdf *W, const double *y, double *A, double *B, double *G2)
{
  if (W->Case == OTD_COMMONEDGE) {
    double y1 = y[0], y12 = y1*y1;
    double A2 = W->A2, BP2 = W->BP2, L2 = W->L2, AdBP = W->AdBP, AdL = W->AdL, BPdL = W->BPdL;
    pqr_abg2((BP2 - 2.0*BPdL + L2)*y12,
             y1*(L2 + BPdL*(y1-1.0) - (AdL + L2 - AdBP)*y1),
             L2 - 2*AdL*y1 - 2*L2*y1 + A2*y12 + 2*AdL*y12 + L2*y12, A+0, B+0, G2+0);
    pqr_abg2(BP2*y12,
             (BPdL + AdBP*y1 - BPdL*y1)*y1,
             L2 + 2*AdL*y1 - 2*L2*y1 + (A2 - 2*AdL + L2)*y12, A+1, B+1, G2+1);
    pqr_abg2((A2 + 2*AdBP + BP2)*y12,
             y1*(AdL*(-1.0 + y1) + BPdL*(-1.0 + y1) - (AdBP + BP2)*y1),
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

/* the four quadratic-form integrals p=-3,-1,1,2 and their y-moments
   (QIFIPPITaylorDuffyV2P0.cc:430-478) */
static void tdf_qfp(double P, double Q2, double Iq[4], double Iy[4])
{
  double P2 = P*P, PP1 = P + 1.0;
  if (fabs(Q2) < 1.0e-8) {
    double SignP = P > 0.0 ? 1.0 : -1.0;
    Iq[0] = SignP*(P + 0.5)/(P2*PP1*PP1);
    Iy[0] = -P*Iq[0] + SignP/(P*PP1);
    Iq[1] = SignP*log(1.0 + 1.0/P);
    Iy[1] = -P*Iq[1] + SignP*1.0;
    Iq[2] = SignP*(P + 0.5);
    Iy[2] = -P*Iq[2] + SignP*(P*PP1 + 1.0/3.0);
    Iq[3] = P*PP1 + 1.0/3.0;
    Iy[3] = -P*Iq[3] + P*P2 + (3.0/2.0)*P2 + P + 0.25;
  } else {
    double S2 = P*P + Q2, S = sqrt(S2), S3 = S*S2, S4 = S2*S2;
    double T2 = PP1*PP1 + Q2, T = sqrt(T2), T3 = T*T2, T4 = T2*T2;
    double LogFac = log((PP1 + T)/(P + S));
    Iq[0] = (PP1/T - P/S)/Q2;
    Iy[0] = -P*Iq[0] + 1.0/S - 1.0/T;
    Iq[1] = LogFac;
    Iy[1] = -P*Iq[1] + (T - S);
    Iq[2] = (T + P*(T - S) + Q2*LogFac)/2.0;
    Iy[2] = -P*Iq[2] + (T3 - S3)/3.0;
    Iq[3] = (T2 + S2)/2.0 - 1.0/6.0;
    Iy[3] = -P*Iq[3] + (T4 - S4)/4.0;
  }
}

/* Polynomial coefficients of the Duffy-transformed numerators,
   common-edge case (QIFIPPITaylorDuffyV2P0.cc:484-757).
   PCross[slot][region][n][j]: slot 0..2 x-x' components, 3..5 x cross x'
   components; n = power of the radial variable; j=0 constant, j=1
   coefficient of the innermost variable. */
static void tdf_script_p_ce(tdf *W, double y1)
{
  double (*PC)[6][4][2] = W->PCross, (*PX)[6][4][2] = W->PXXEE;
  const double *A = W->A, *B = W->B, *AP = W->AP, *BP = W->BP;
  const double *V0xAP = W->V0xAP, *V0xBP = W->V0xBP, *AxV0 = W->AxV0, *BxV0 = W->BxV0;
  const double *AxAP = W->AxAP, *AxBP = W->AxBP, *BxAP = W->BxAP, *BxBP = W->BxBP;
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

    double *q;
    q = PC[m+3][0][0]; (void)q;
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

  /* entries not assigned here stay zero from the caller's clear */
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

/* common-vertex case (QIFIPPITaylorDuffyV2P0.cc:762-829): regions
   d=0 (x-triangle outermost) and d=1 (x'-triangle outermost). */
static void tdf_script_p_cv(tdf *W, double y1, double y2)
{
  double (*PC)[6][4][2] = W->PCross, (*PX)[6][4][2] = W->PXXEE;
  const double *A = W->A, *B = W->B, *AP = W->AP, *BP = W->BP;
  const double *V0xAP = W->V0xAP, *V0xBP = W->V0xBP, *AxV0 = W->AxV0, *BxV0 = W->BxV0;
  const double *AxAP = W->AxAP, *AxBP = W->AxBP, *BxAP = W->BxAP, *BxBP = W->BxBP;
  for (int m = 0; m < 3; m++) {
    PC[m][0][1][0] = -AP[m]*y2 + B[m]*y1 + A[m];
    PC[m][0][1][1] = -BP[m]*y2;
    PC[m][0][2][0] = 0.0; PC[m][0][2][1] = 0.0;
    PC[m][1][1][0] = A[m]*y2 - BP[m]*y1 - AP[m];
    PC[m][1][1][1] = B[m]*y2;
    PC[m][1][2][0] = 0.0; PC[m][1][2][1] = 0.0;
    PC[m+3][0][1][0] = V0xAP[m]*y2 + BxV0[m]*y1 + AxV0[m];
    PC[m+3][0][1][1] = V0xBP[m]*y2;
    PC[m+3][0][2][0] = (BxAP[m]*y1 + AxAP[m])*y2;
    PC[m+3][0][2][1] = (BxBP[m]*y1 + AxBP[m])*y2;
    PC[m+3][1][1][0] = AxV0[m]*y2 + V0xBP[m]*y1 + V0xAP[m];
    PC[m+3][1][1][1] = BxV0[m]*y2;
    PC[m+3][1][2][0] = (AxBP[m]*y1 + AxAP[m])*y2;
    PC[m+3][1][2][1] = (BxBP[m]*y1 + BxAP[m])*y2;
  }
  PX[0][0][0][0] = 1.0;  PX[0][1][0][0] = 1.0;
  PX[1][0][1][0] = y2;   PX[1][1][1][0] = 1.0;
  PX[2][0][1][1] = y2;   PX[2][1][1][0] = y1;
  PX[3][0][1][0] = 1.0;  PX[3][1][1][0] = y2;
  PX[4][0][2][0] = y2;   PX[4][1][2][0] = y2;
  PX[5][0][2][1] = y2;   PX[5][1][2][0] = y1*y2;
  PX[6][0][1][0] = y1;   PX[6][1][1][1] = y2;
  PX[7][0][2][0] = y1*y2; PX[7][1][2][1] = y2;
  PX[8][0][2][1] = y1*y2; PX[8][1][2][1] = y1*y2;
}

/* integrand (QIFIPPITaylorDuffyV2P0.cc:91-221) */
static void tdf_sum(unsigned ndim, const double *y, void *parms, unsigned nfun, double *f)
{
  (void)ndim; (void)nfun;
  tdf *W = (tdf *)parms;
  W->nCalls++;
  int NumRegions, nOffset, nMinC, nMaxC, nMinX, nMaxX;
  double Jac;
  if (W->Case == OTD_COMMONEDGE) { NumRegions = 6; nOffset = 2; Jac = y[0]; nMinC = 1; nMaxC = 3; nMinX = 0; nMaxX = 3; }
  else { NumRegions = 2; nOffset = 3; Jac = y[1]; nMinC = 1; nMaxC = 2; nMinX = 0; nMaxX = 2; }
  double A[6], B[6], G2[6];
  tdf_abg(W, y, A, B, G2);
  double JrM3[6][6], LrM3[6][6], Jrp[3][6][6], Lrp[3][6][6];
  double Iq[4], Iy[4];
  for (int d = 0; d < NumRegions; d++) {
    tdf_qfp(B[d], G2[d], Iq, Iy);
    double A2 = A[d]*A[d], A3 = A2*A[d];
    for (int n = nMinC + nOffset; n <= nMaxC + nOffset; n++) {
      JrM3[d][n] = Iq[0]/(A3*(1.0 + n - 3.0));
      LrM3[d][n] = Iy[0]/(A3*(1.0 + n - 3.0));
    }
    for (int n = nMinX + nOffset; n <= nMaxX + nOffset; n++) {
      Jrp[0][d][n] = Iq[1]/(A[d]*(1.0 + n - 1.0));
      Lrp[0][d][n] = Iy[1]/(A[d]*(1.0 + n - 1.0));
      Jrp[1][d][n] = A[d]*Iq[2]/(1.0 + n + 1.0);
      Lrp[1][d][n] = A[d]*Iy[2]/(1.0 + n + 1.0);
      Jrp[2][d][n] = A2*Iq[3]/(1.0 + n + 2.0);
      Lrp[2][d][n] = A2*Iy[3]/(1.0 + n + 2.0);
    }
  }
  if (W->Case == OTD_COMMONEDGE) tdf_script_p_ce(W, y[0]);
  else tdf_script_p_cv(W, y[0], y[1]);
  memset(f, 0, 33*sizeof(double));
  int ns = 0;
  for (int np = 0; np < 6; np++, ns++)
    for (int d = 0; d < NumRegions; d++)
      for (int n = nMinC; n <= nMaxC; n++)
        f[ns] += Jac*(W->PCross[np][d][n][0]*JrM3[d][n+nOffset] + W->PCross[np][d][n][1]*LrM3[d][n+nOffset]);
  for (int nk = 0; nk < 3; nk++)
    for (int np = 0; np < 9; np++, ns++)
      for (int d = 0; d < NumRegions; d++)
        for (int n = nMinX; n <= nMaxX; n++)
          f[ns] += Jac*(W->PXXEE[np][d][n][0]*Jrp[nk][d][n+nOffset] + W->PXXEE[np][d][n][1]*Lrp[nk][d][n+nOffset]);
}

/* ComputeQIFIPPIData_TaylorDuffyV2P0 (QIFIPPITaylorDuffyV2P0.cc:226-312) */
static void qifippi_td(double *V1, double *V2, double *V3, double *V2P, double *V3P,
                       OQIFIPPIData *Q, size_t *nevals)
{
  tdf *W = (tdf *)calloc(1, sizeof(tdf));
  double A[3], AP[3], B[3], BP[3], L[3];
  sub3(V2, V1, A); sub3(V3, V2, B); sub3(V2P, V1, AP); sub3(V3P, V2P, BP); sub3(BP, B, L);
  W->A2 = dot3(A, A); W->B2 = dot3(B, B); W->AP2 = dot3(AP, AP); W->BP2 = dot3(BP, BP);
  W->L2 = dot3(L, L); W->AdB = dot3(A, B); W->AdAP = dot3(A, AP); W->AdBP = dot3(A, BP);
  W->AdL = dot3(A, L); W->BdAP = dot3(B, AP); W->BdBP = dot3(B, BP); W->APdBP = dot3(AP, BP);
  W->BPdL = dot3(BP, L);
  memcpy(W->A, A, sizeof(A)); memcpy(W->B, B, sizeof(B));
  memcpy(W->AP, AP, sizeof(AP)); memcpy(W->BP, BP, sizeof(BP));
  cross3(V1, AP, W->V0xAP); cross3(V1, BP, W->V0xBP);
  cross3(A, V1, W->AxV0); cross3(B, V1, W->BxV0);
  cross3(A, AP, W->AxAP); cross3(B, AP, W->BxAP);
  cross3(A, BP, W->AxBP); cross3(B, BP, W->BxBP);
  /* common edge iff V2 and V2P agree in single precision (Vector.cc:279-284) */
  int ce = ((float)V2[0] == (float)V2P[0]) && ((float)V2[1] == (float)V2P[1]) && ((float)V2[2] == (float)V2P[2]);
  W->Case = ce ? OTD_COMMONEDGE : OTD_COMMONVERTEX;
  double R[33], E[33];
  const double lo[2] = {0, 0}, hi[2] = {1, 1};
  o_hcubature(33, tdf_sum, W, ce ? 1 : 2, lo, hi, 10000, 0.0, 1.0e-5, R, E, NULL);
  if (nevals) *nevals = W->nCalls;
  memcpy(Q->xMxpRM3, R+0, 3*sizeof(double));
  memcpy(Q->xXxpRM3, R+3, 3*sizeof(double));
  memcpy(Q->uvupvpRM1, R+6, 9*sizeof(double));
  memcpy(Q->uvupvpR1, R+15, 9*sizeof(double));
  memcpy(Q->uvupvpR2, R+24, 9*sizeof(double));
  free(W);
}

/* ComputeQIFIPPIData (QIFIPPICubature.cc:465-487) */
void o_compute_qifippi(double **Va, double **Vb, int ncv, OQIFIPPIData *Q, size_t *nevals)
{
  if (ncv == 0) qifippi_cubature(Va, Vb, Q, nevals);
  else qifippi_td(Va[0], Va[1], Va[2], Vb[1], Vb[2], Q, nevals);
  for (int m = 0; m < 3; m++) Q->x0[m] = (Va[0][m] + Va[1][m] + Va[2][m])/3.0;
}

/* ------------------------------------------------------------------ */
/* cache: open-addressing table keyed on the 60-byte float key          */
/* ------------------------------------------------------------------ */
typedef struct { float key[15]; OQIFIPPIData *val; } fentry;
struct OFIPPICache { size_t cap, n, hits; fentry *tab; };

OFIPPICache *o_fippi_cache_new(void)
{
  OFIPPICache *C = (OFIPPICache *)calloc(1, sizeof(OFIPPICache));
  C->cap = 1024;
  C->tab = (fentry *)calloc(C->cap, sizeof(fentry));
  return C;
}

void o_fippi_cache_free(OFIPPICache *C)
{
  if (!C) return;
  for (size_t i = 0; i < C->cap; i++) free(C->tab[i].val);
  free(C->tab); free(C);
}

size_t o_fippi_cache_size(OFIPPICache *C) { return C->n; }
size_t o_fippi_cache_hits(OFIPPICache *C) { return C->hits; }

static uint64_t key_hash(const float *k)
{
  const unsigned char *p = (const unsigned char *)k;
  uint64_t h = 1469598103934665603ULL;
  for (int i = 0; i < 60; i++) { h ^= p[i]; h *= 1099511628211ULL; }
  return h;
}

static void cache_grow(OFIPPICache *C)
{
  size_t ncap = C->cap*2;
  fentry *nt = (fentry *)calloc(ncap, sizeof(fentry));
  for (size_t i = 0; i < C->cap; i++)
    if (C->tab[i].val) {
      size_t h = key_hash(C->tab[i].key) & (ncap-1);
      while (nt[h].val) h = (h+1) & (ncap-1);
      nt[h] = C->tab[i];
    }
  free(C->tab); C->tab = nt; C->cap = ncap;
}

/* FIPPICache::GetQIFIPPIData (FIPPICache.cc:135-187) */
OQIFIPPIData *o_fippi_cache_get(OFIPPICache *C, double **OVa, double **OVb, int ncv)
{
  float K[15];
  const double *src[5] = {OVa[1], OVa[2], OVb[0], OVb[1], OVb[2]};
  for (int s = 0; s < 5; s++)
    for (int m = 0; m < 3; m++)
      K[3*s+m] = ((float)src[s][m]) - ((float)OVa[0][m]);
  size_t h = key_hash(K) & (C->cap-1);
  while (C->tab[h].val) {
    if (!memcmp(C->tab[h].key, K, sizeof(K))) { C->hits++; return C->tab[h].val; }
    h = (h+1) & (C->cap-1);
  }
  OQIFIPPIData *Q = (OQIFIPPIData *)malloc(sizeof(OQIFIPPIData));
  o_compute_qifippi(OVa, OVb, ncv, Q, NULL);
  memcpy(C->tab[h].key, K, sizeof(K));
  C->tab[h].val = Q;
  if (++C->n*2 > C->cap) cache_grow(C);
  return Q;
}
