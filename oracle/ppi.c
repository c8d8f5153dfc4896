/*
 * ppi.c -- oracle: panel-panel integrals.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates libs/libscuff/PanelPanelInteractions.cc:84-561 (ExpRel, the
 * inner cubature integrand, fixed-rule cubature, the algorithm
 * dispatch), AssessPanelPair.cc:52-172,250-455 (common-vertex detection
 * and canonical ordering) and QDFIPPI.cc:67-260 (recombination of the
 * Q-independent FIPPIs into the Q-dependent singular terms).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static void sub3(const double *a, const double *b, double *c) { c[0]=a[0]-b[0]; c[1]=a[1]-b[1]; c[2]=a[2]-b[2]; }
static double *cross3(const double *a, const double *b, double *c)
{ c[0]=a[1]*b[2]-a[2]*b[1]; c[1]=a[2]*b[0]-a[0]*b[2]; c[2]=a[0]*b[1]-a[1]*b[0]; return c; }
static double norm3(const double *a) { return sqrt(dot3(a, a)); }
static double dist3(const double *a, const double *b)
{ double d = 0; for (int i = 0; i < 3; i++) d += (a[i]-b[i])*(a[i]-b[i]); return sqrt(d); }

/* ExpRel(x,n) = sum_{m>=n} x^m/m!, truncated at |term|<1e-8|sum|
   (PanelPanelInteractions.cc:84-103) */
static zdouble exprel(zdouble x, int n)
{
  int m;
  zdouble Term, Sum;
  for (Term = 1.0, m = 1; m < n; m++) Term *= x/((double)m);
  for (Sum = 0.0; m < 100; m++) {
    Term *= x/((double)m);
    Sum += Term;
    double t2 = creal(Term)*creal(Term) + cimag(Term)*cimag(Term);
    double s2 = creal(Sum)*creal(Sum) + cimag(Sum)*cimag(Sum);
    if (t2 < 1.0e-8*1.0e-8*s2) break;
  }
  return Sum;
}

/* AssembleInnerPPIIntegrand, non-periodic branch (:108-201), with the
   displacement gradients (:176-180) and the angular derivatives about the
   axes with rotation generators Gamma (:182-199) */
static void inner_integrand(double wp, const double *R, const double *X, const double *F, const double *FP,
                            zdouble k, int DeSing, zdouble *HI, zdouble *GHI,
                            int ntq, const double *Gamma, zdouble *DTI)
{
  zdouble ik = I*k, ik2 = ik*ik;
  zdouble hPlus = dot3(F, FP) + 4.0/ik2;
  double FxFP[3];
  cross3(F, FP, FxFP);
  double r2 = dot3(R, R), r = sqrt(r2);
  zdouble Phi;
  if (DeSing) Phi = exprel(ik*r, 4)/(4.0*M_PI*r);
  else Phi = cexp(ik*r)/(4.0*M_PI*r);
  if (!isfinite(creal(Phi))) Phi = 0.0;
  Phi *= wp;
  zdouble Psi = Phi*(ik - 1.0/r)/r;
  zdouble Zeta = Phi*(ik2 - 3.0*ik/r + 3.0/r2)/r2;
  double hTimes = dot3(FxFP, R);
  HI[0] += hPlus*Phi;
  HI[1] += hTimes*Psi;
  if (GHI)
    for (int m = 0; m < 3; m++) {
      GHI[2*m+0] += R[m]*hPlus*Psi;
      GHI[2*m+1] += R[m]*hTimes*Zeta + FxFP[m]*Psi;
    }
  for (int t = 0; t < ntq; t++) {
    double dX[3] = {0, 0, 0}, dF[3] = {0, 0, 0}, dFxFP[3];
    for (int m = 0; m < 3; m++)
      for (int n = 0; n < 3; n++) {
        dX[m] += Gamma[9*t + m + 3*n]*X[n];
        dF[m] += Gamma[9*t + m + 3*n]*F[n];
      }
    double Puv = dot3(R, dX);
    cross3(dF, FP, dFxFP);
    DTI[2*t+0] += hPlus*Puv*Psi + dot3(dF, FP)*Phi;
    DTI[2*t+1] += hTimes*Puv*Zeta + (dot3(dFxFP, R) + dot3(FxFP, dX))*Psi;
  }
}

/* GetPPIs_Cubature (:211-350) */
static void ppis_cubature(OPPIArgs *A, int DeSing, int HighOrder,
                          double **Va, const double *Qa, double **Vb, const double *Qb)
{
  double *V0 = Va[0], Av[3], Bv[3], *V0P = Vb[0], APv[3], BPv[3];
  sub3(Va[1], Va[0], Av); sub3(Va[2], Va[0], Bv);
  sub3(Vb[1], Vb[0], APv); sub3(Vb[2], Vb[0], BPv);
  int want_grad = A->NumGradientComponents > 0;
  int ntq = (A->GammaMatrix && A->NumTorqueAxes > 0) ? A->NumTorqueAxes : 0;
  int np;
  const double *TCR = o_tcr(HighOrder ? 20 : 4, &np);
  zdouble H[2] = {0, 0}, G[6] = {0, 0, 0, 0, 0, 0}, T[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < np; i++) {
    double u = TCR[3*i], v = TCR[3*i+1], w = TCR[3*i+2];
    double X[3], F[3];
    for (int m = 0; m < 3; m++) { X[m] = V0[m] + u*Av[m] + v*Bv[m]; F[m] = X[m] - Qa[m]; }
    zdouble HI[2] = {0, 0}, GI[6] = {0, 0, 0, 0, 0, 0}, TI[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < np; j++) {
      double up = TCR[3*j], vp = TCR[3*j+1], wp = TCR[3*j+2];
      double XP[3], FP[3], R[3];
      for (int m = 0; m < 3; m++) {
        XP[m] = V0P[m] + up*APv[m] + vp*BPv[m];
        FP[m] = XP[m] - Qb[m];
        R[m] = X[m] - XP[m];
      }
      inner_integrand(wp, R, X, F, FP, A->k, DeSing, HI, want_grad ? GI : NULL, ntq, A->GammaMatrix, TI);
    }
    H[0] += w*HI[0];
    H[1] += w*HI[1];
    if (want_grad) for (int m = 0; m < 6; m++) G[m] += w*GI[m];
    for (int m = 0; m < 2*ntq; m++) T[m] += w*TI[m];
  }
  A->H[0] = H[0]; A->H[1] = H[1];
  if (want_grad) memcpy(A->GradH, G, sizeof(G));
  if (ntq) memcpy(A->dHdT, T, sizeof(T));
}

/* ------------------------------------------------------------------ */
/* common vertices (AssessPanelPair.cc:52-172)                          */
/* ------------------------------------------------------------------ */
static int assess_vertices(double **Va, double **Vb, double Radius)
{
  int ncv = 0, CVIa[3], CVIb[3];
  double thr = fabs(1.0e-6*Radius);
  for (int ia = 0; ia < 3; ia++)
    for (int ib = 0; ib < 3; ib++)
      if (fabs(Va[ia][0]-Vb[ib][0]) + fabs(Va[ia][1]-Vb[ib][1]) + fabs(Va[ia][2]-Vb[ib][2]) <= 3*thr) {
        CVIa[ncv] = ia; CVIb[ncv] = ib; ncv++;
      }
  if (ncv == 0) return 0;
  double *OVa[3] = {Va[0], Va[1], Va[2]}, *OVb[3] = {Vb[0], Vb[1], Vb[2]};
  if (ncv == 1) {
    Va[0] = OVa[CVIa[0]]; Va[1] = OVa[(CVIa[0]+1)%3]; Va[2] = OVa[(CVIa[0]+2)%3];
    Vb[0] = OVb[CVIb[0]]; Vb[1] = OVb[(CVIb[0]+1)%3]; Vb[2] = OVb[(CVIb[0]+2)%3];
  } else if (ncv == 2) {
    Va[0] = OVa[CVIa[0]]; Va[1] = OVa[CVIa[1]]; Va[2] = OVa[3-CVIa[0]-CVIa[1]];
    Vb[0] = OVb[CVIb[0]]; Vb[1] = OVb[CVIb[1]]; Vb[2] = OVb[3-CVIb[0]-CVIb[1]];
  } else if (ncv == 3) {
    Va[0] = OVa[CVIa[0]]; Va[1] = OVa[CVIa[1]]; Va[2] = OVa[CVIa[2]];
    Vb[0] = OVb[CVIb[0]]; Vb[1] = OVb[CVIb[1]]; Vb[2] = OVb[CVIb[2]];
  }
  return ncv;
}

int o_assess_panel_pair(OSurface *Sa, int npa, OSurface *Sb, int npb,
                        double *rRel, double **Va, double **Vb)
{
  OPanel *Pa = Sa->Panels + npa, *Pb = Sb->Panels + npb;
  for (int i = 0; i < 3; i++) { Va[i] = Sa->Vertices + 3*Pa->VI[i]; Vb[i] = Sb->Vertices + 3*Pb->VI[i]; }
  double DC, rRel2, rMax = fmax(Pa->Radius, Pb->Radius);
  DC = Pa->Centroid[0] - Pb->Centroid[0]; rRel2 = DC*DC;
  DC = Pa->Centroid[1] - Pb->Centroid[1]; rRel2 += DC*DC;
  DC = Pa->Centroid[2] - Pb->Centroid[2]; rRel2 += DC*DC;
  *rRel = sqrt(rRel2)/rMax;
  if (*rRel > 2.0) return 0;
  return assess_vertices(Va, Vb, rMax);
}

/* vertex "less than" with 1e-6 relative coordinate tolerance (:250-267) */
static int vlt(const double *V1, const double *V2)
{
  double DV;
  DV = V1[0] - V2[0]; if (fabs(DV) > 1.0e-6*fabs(V1[0])) return DV < 0.0 ? 1 : 0;
  DV = V1[1] - V2[1]; if (fabs(DV) > 1.0e-6*fabs(V1[1])) return DV < 0.0 ? 1 : 0;
  DV = V1[2] - V2[2]; if (fabs(DV) > 1.0e-6*fabs(V1[2])) return DV < 0.0 ? 1 : 0;
  return 1;
}

/* CanonicallyOrderVertices (:342-455), including its ncv==0 branch,
   which compares Va[iMina] with Va[iMinb] and fills OVb from Vb using
   panel-a's sort indices. */
int o_canonically_order(double **Va, double **Vb, int ncv, double **OVa, double **OVb)
{
  int iMina, iMeda = 0, iMaxa, iMinb, iMaxb;
  if (ncv == 3) {
    if (vlt(Va[0], Va[1])) { iMina = 0; iMaxa = 1; } else { iMina = 1; iMaxa = 0; }
    if (vlt(Va[2], Va[iMina])) { iMeda = iMina; iMina = 2; }
    else if (vlt(Va[2], Va[iMaxa])) iMeda = 2;
    else { iMeda = iMaxa; iMaxa = 2; }
    OVa[0] = Va[iMina]; OVa[1] = Va[iMeda]; OVa[2] = Va[iMaxa];
    OVb[0] = Vb[iMina]; OVb[1] = Vb[iMeda]; OVb[2] = Vb[iMaxa];
    return 0;
  } else if (ncv == 2) {
    int a01Min = vlt(Va[0], Va[1]) ? 0 : 1;
    if (vlt(Va[2], Vb[2])) {
      OVa[0] = Va[a01Min]; OVa[1] = Va[1-a01Min]; OVa[2] = Va[2];
      OVb[0] = Vb[a01Min]; OVb[1] = Vb[1-a01Min]; OVb[2] = Vb[2];
      return 0;
    }
    OVa[0] = Vb[a01Min]; OVa[1] = Vb[1-a01Min]; OVa[2] = Vb[2];
    OVb[0] = Va[a01Min]; OVb[1] = Va[1-a01Min]; OVb[2] = Va[2];
    return 1;
  } else if (ncv == 1) {
    int a12Min = vlt(Va[1], Va[2]) ? 1 : 2;
    int b12Min = vlt(Vb[1], Vb[2]) ? 1 : 2;
    if (vlt(Va[a12Min], Vb[b12Min])) {
      OVa[0] = Va[0]; OVa[1] = Va[a12Min]; OVa[2] = Va[3-a12Min];
      OVb[0] = Vb[0]; OVb[1] = Vb[b12Min]; OVb[2] = Vb[3-b12Min];
      return 0;
    }
    OVa[0] = Vb[0]; OVa[1] = Vb[b12Min]; OVa[2] = Vb[3-b12Min];
    OVb[0] = Va[0]; OVb[1] = Va[a12Min]; OVb[2] = Va[3-a12Min];
    return 1;
  }
  if (vlt(Va[0], Va[1])) { iMina = 0; iMaxa = 1; } else { iMina = 1; iMaxa = 0; }
  if (vlt(Va[2], Va[iMina])) { iMeda = iMina; iMina = 2; }
  else if (vlt(Va[2], Va[iMaxa])) iMeda = 2;
  else { iMeda = iMaxa; iMaxa = 2; }
  if (vlt(Vb[0], Vb[1])) { iMinb = 0; iMaxb = 1; } else { iMinb = 1; iMaxb = 0; }
  if (vlt(Vb[2], Vb[iMinb])) iMinb = 2;
  else if (vlt(Vb[2], Vb[iMaxb])) { }
  else iMaxb = 2;
  (void)iMaxb;
  if (vlt(Va[iMina], Va[iMinb])) {
    OVa[0] = Va[iMina]; OVa[1] = Va[iMeda]; OVa[2] = Va[iMaxa];
    OVb[0] = Vb[iMina]; OVb[1] = Vb[iMeda]; OVb[2] = Vb[iMaxa];
    return 0;
  }
  OVb[0] = Va[iMina]; OVb[1] = Va[iMeda]; OVb[2] = Va[iMaxa];
  OVa[0] = Vb[iMina]; OVa[1] = Vb[iMeda]; OVa[2] = Vb[iMaxa];
  return 1;
}

/* ------------------------------------------------------------------ */
/* GetQDFIPPIData (QDFIPPI.cc:67-260)                                  */
/* ------------------------------------------------------------------ */
enum { Q_ONE = 0, Q_UP, Q_VP, Q_U, Q_UUP, Q_UVP, Q_V, Q_VUP, Q_VVP };
static const double r0mom[9] = {1.0/4.0, 1.0/6.0, 1.0/12.0, 1.0/6.0, 1.0/9.0,
                                 1.0/18.0, 1.0/12.0, 1.0/18.0, 1.0/36.0};

static void qdfippi(double **Va, double *Qa, double **Vb, double *Qb, int ncv,
                    OFIPPICache *FC, OQDFIPPIData *D)
{
  OQIFIPPIData local, *Q;
  double *OVa[3], *OQa, *OVb[3], *OQb;
  int Flipped;
  if (FC) {
    Flipped = o_canonically_order(Va, Vb, ncv, OVa, OVb);
    if (Flipped) { OQa = Qb; OQb = Qa; } else { OQa = Qa; OQb = Qb; }
    Q = o_fippi_cache_get(FC, OVa, OVb, ncv);
  } else {
    Flipped = 0;
    for (int i = 0; i < 3; i++) { OVa[i] = Va[i]; OVb[i] = Vb[i]; }
    OQa = Qa; OQb = Qb;
    o_compute_qifippi(OVa, OVb, ncv, &local, NULL);
    Q = &local;
  }
  double A[3], B[3], V0mQ[3], AP[3], BP[3], V0PmQP[3];
  sub3(OVa[1], OVa[0], A); sub3(OVa[2], OVa[1], B); sub3(OVa[0], OQa, V0mQ);
  sub3(OVb[1], OVb[0], AP); sub3(OVb[2], OVb[1], BP); sub3(OVb[0], OQb, V0PmQP);
  double VdVP = dot3(V0mQ, V0PmQP), VdAP = dot3(V0mQ, AP), VdBP = dot3(V0mQ, BP);
  double AdVP = dot3(A, V0PmQP), AdAP = dot3(A, AP), AdBP = dot3(A, BP);
  double BdVP = dot3(B, V0PmQP), BdAP = dot3(B, AP), BdBP = dot3(B, BP);
  double QamQb[3], QaxQb[3], V0mV0P[3], Scr[3];
  double *V0 = OVa[0], *V0P = OVb[0];
  sub3(V0, V0P, V0mV0P);
  sub3(Qa, Qb, QamQb);
  cross3(Qa, Qb, QaxQb);
  double ht0000 = dot3(QaxQb, V0mV0P) + dot3(QamQb, cross3(V0, V0P, Scr));
  double ht1000 = dot3(QaxQb, A) + dot3(QamQb, cross3(A, V0P, Scr));
  double ht0100 = dot3(QaxQb, B) + dot3(QamQb, cross3(B, V0P, Scr));
  double ht0010 = -dot3(QaxQb, AP) + dot3(QamQb, cross3(V0, AP, Scr));
  double ht0001 = -dot3(QaxQb, BP) + dot3(QamQb, cross3(V0, BP, Scr));
  double ht1010 = dot3(QamQb, cross3(A, AP, Scr));
  double ht1001 = dot3(QamQb, cross3(A, BP, Scr));
  double ht0110 = dot3(QamQb, cross3(B, AP, Scr));
  double ht0101 = dot3(QamQb, cross3(B, BP, Scr));

  D->hTimesRM3 = dot3(QaxQb, Q->xMxpRM3) + dot3(QamQb, Q->xXxpRM3);
  {
    /* translation correction applied when the cached record was
       computed for a translated copy of this pair (QDFIPPI.cc:143-158) */
    double L[3];
    double **Vc = Flipped ? Vb : Va;
    for (int m = 0; m < 3; m++) L[m] = Q->x0[m] - (Vc[0][m] + Vc[1][m] + Vc[2][m])/3.0;
    if (norm3(L) > 1.0e-8*norm3(A))
      D->hTimesRM3 -= dot3(cross3(L, QamQb, Scr), Q->xMxpRM3);
  }
  const double *RM1 = Q->uvupvpRM1, *R1 = Q->uvupvpR1, *R2 = Q->uvupvpR2;
#define HDOT(M) (M[Q_ONE]*VdVP + M[Q_UP]*VdAP + M[Q_VP]*VdBP + M[Q_U]*AdVP + M[Q_UUP]*AdAP \
                 + M[Q_UVP]*AdBP + M[Q_V]*BdVP + M[Q_VUP]*BdAP + M[Q_VVP]*BdBP)
#define HTIMES(M) (M[Q_ONE]*ht0000 + M[Q_U]*ht1000 + M[Q_V]*ht0100 + M[Q_UP]*ht0010 \
                   + M[Q_VP]*ht0001 + M[Q_UUP]*ht1010 + M[Q_UVP]*ht1001 + M[Q_VUP]*ht0110 \
                   + M[Q_VVP]*ht0101)
  D->hDotRM1 = HDOT(RM1);
  D->hNablaRM1 = 4.0*RM1[Q_ONE];
  D->hTimesRM1 = HTIMES(RM1);
  D->hDotR0 = HDOT(r0mom);
  D->hNablaR0 = 1.0;
  D->hTimesR0 = HTIMES(r0mom);
  D->hDotR1 = HDOT(R1);
  D->hNablaR1 = 4.0*R1[Q_ONE];
  D->hTimesR1 = HTIMES(R1);
  D->hDotR2 = HDOT(R2);
  D->hNablaR2 = 4.0*R2[Q_ONE];
#undef HDOT
#undef HTIMES
  if (Flipped) {
    D->hTimesRM3 *= -1.0; D->hTimesRM1 *= -1.0; D->hTimesR0 *= -1.0; D->hTimesR1 *= -1.0;
  }
}

void o_ppi_init(OPPIArgs *A)
{
  memset(A, 0, sizeof(*A));
}

/* GetPanelPanelInteractions (:357-561), compact (non-periodic), no
   displacement, no torque. */
void o_get_ppis(OPPIArgs *A)
{
  OSurface *Sa = A->Sa, *Sb = A->Sb;
  OPanel *Pa = Sa->Panels + A->npa, *Pb = Sb->Panels + A->npb;
  double *Qa = Sa->Vertices + 3*Pa->VI[A->iQa];
  double *Qb = Sb->Vertices + 3*Pb->VI[A->iQb];
  double *Va[3], *Vb[3], rRel;
  zdouble k = A->k;
  int ncv = o_assess_panel_pair(Sa, A->npa, Sb, A->npb, &rRel, Va, Vb);

  if (rRel > 4.0) {
    A->WhichAlgorithm = OPPI_LOCUBATURE;
    ppis_cubature(A, 0, 0, Va, Qa, Vb, Qb);
    return;
  }
  double kR = cabs(k*fmax(Pa->Radius, Pb->Radius));
  int SW = kR > 1.0*M_PI, VSW = kR > 20.0;
  if (SW && ncv == 0) {
    A->WhichAlgorithm = OPPI_HOCUBATURE;
    ppis_cubature(A, 0, 1, Va, Qa, Vb, Qb);
    return;
  }
  if (ncv == 3 || (ncv > 0 && (SW || A->ForceTaylorDuffy))) {
    int PIndex[3] = {OTD_UNITY, OTD_PMCHWG1, OTD_PMCHWC};
    int KIndex[3] = {OTD_HELMHOLTZ, OTD_HELMHOLTZ, OTD_GRADHELMHOLTZ};
    zdouble KParam[3] = {k, k, k}, Result[3], Error[3];
    OTDArgs T;
    o_taylor_duffy_init(&T);
    T.WhichCase = ncv;
    T.NumPKs = (ncv == 3) ? 2 : 3;
    T.PIndex = PIndex; T.KIndex = KIndex; T.KParam = KParam;
    T.V1 = Va[0]; T.V2 = Va[1]; T.V3 = Va[2]; T.V2P = Vb[1]; T.V3P = Vb[2];
    T.Q = Qa; T.QP = Qb;
    T.Result = Result; T.Error = Error;
    if (VSW) {
      A->WhichAlgorithm = OPPI_HKTD;
      KIndex[0] = OTD_HIGHK_HELMHOLTZ; KIndex[1] = OTD_HIGHK_HELMHOLTZ; KIndex[2] = OTD_HIGHK_GRADHELMHOLTZ;
    } else A->WhichAlgorithm = OPPI_TD;
    o_taylor_duffy(&T);
    A->H[0] = Result[1] - 4.0*Result[0]/(k*k);
    A->H[1] = (ncv == 3) ? 0.0 : Result[2];
    memset(A->GradH, 0, sizeof(A->GradH));
    memset(A->dHdT, 0, sizeof(A->dHdT));
    return;
  }
  zdouble GradSave[6], TorqueSave[6];
  A->WhichAlgorithm = OPPI_DESING;
  int ntq = A->NumTorqueAxes;
  if (A->NumGradientComponents > 0 || (ntq > 0 && A->GammaMatrix)) {
    ppis_cubature(A, 0, 1, Va, Qa, Vb, Qb);
    memcpy(GradSave, A->GradH, sizeof(GradSave));
    memcpy(TorqueSave, A->dHdT, sizeof(TorqueSave));
  }
  int ng = A->NumGradientComponents;
  A->NumGradientComponents = 0;
  A->NumTorqueAxes = 0;
  ppis_cubature(A, 1, 0, Va, Qa, Vb, Qb);
  A->NumGradientComponents = ng;
  A->NumTorqueAxes = ntq;
  OQDFIPPIData D;
  qdfippi(Va, Qa, Vb, Qb, ncv, A->FC, &D);
  zdouble ik = I*k, OOIK2 = 1.0/(ik*ik), PF[5];
  PF[0] = 1.0/(4.0*M_PI);
  PF[1] = ik*PF[0]; PF[2] = ik*PF[1]; PF[3] = ik*PF[2]; PF[4] = ik*PF[3];
  A->H[0] += PF[0]*1.0*(D.hDotRM1 + OOIK2*D.hNablaRM1)
           + PF[1]*1.0*(D.hDotR0 + OOIK2*D.hNablaR0)
           + PF[2]*(1.0/2.0)*(D.hDotR1 + OOIK2*D.hNablaR1)
           + PF[3]*(1.0/6.0)*(D.hDotR2 + OOIK2*D.hNablaR2);
  A->H[1] += PF[0]*(-1.0)*D.hTimesRM3
           + PF[2]*(1.0/2.0)*D.hTimesRM1
           + PF[3]*(1.0/3.0)*D.hTimesR0
           + PF[4]*(1.0/6.0)*D.hTimesR1;
  if (ng > 0) memcpy(A->GradH, GradSave, sizeof(GradSave));
  if (ntq > 0 && A->GammaMatrix) memcpy(A->dHdT, TorqueSave, sizeof(TorqueSave));
}
