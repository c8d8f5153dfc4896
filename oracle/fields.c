/*
 * fields.c -- oracle: plane-wave excitation, RHS vector, power and the
 * Mie series.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  o_planewave_fields  PlaneWave::GetFields (libs/libIncField/PlaneWave.cc:71-87)
 *  o_assemble_rhs_pw   AssembleRHSVector with one exterior plane wave
 *                      (libs/libscuff/AssembleRHSVector.cc:147-427): degree-20
 *                      triangle rule on each panel, sources in the exterior
 *                      region enter with a minus sign, E rows divided by ZVAC
 *  o_power_emt         scattered power by energy/momentum transfer from the
 *                      exterior-region matrix (EMTPFT.cc:468-803)
 *  o_power             absorbed power from the crossed overlap integrals
 *                      (OPFT.cc:82-146, :310-364); extinction from the
 *                      inner products of the solution with the RHS
 *  o_mie_sphere        Mie efficiencies from the Bohren-Huffman recurrences
 *                      (logarithmic derivative by downward recursion);
 *                      m_rel = INFINITY-like flag (creal<0) selects the PEC limit
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void o_planewave_fields(const zdouble E0[3], const double nHat[3], zdouble Omega,
                        zdouble Eps, zdouble Mu, const double X[3], zdouble EH[6])
{
  zdouble K = csqrt(Eps*Mu)*Omega;
  zdouble Z = O_ZVAC*csqrt(Mu/Eps);
  zdouble ef = cexp(I*K*(nHat[0]*X[0] + nHat[1]*X[1] + nHat[2]*X[2]));
  EH[0] = E0[0]*ef; EH[1] = E0[1]*ef; EH[2] = E0[2]*ef;
  EH[3] = (nHat[1]*EH[2] - nHat[2]*EH[1])/Z;
  EH[4] = (nHat[2]*EH[0] - nHat[0]*EH[2])/Z;
  EH[5] = (nHat[0]*EH[1] - nHat[1]*EH[0])/Z;
}

/* <f_e, E> and <f_e, H> over one panel with the degree-20 rule
   (TriIntFixed, libs/libTriInt/libTriInt.cc:279-322) */
static void panel_inner(const double *V1, const double *V2, const double *V3,
                        const double *Q, double PreFac, const zdouble E0[3],
                        const double nHat[3], zdouble Omega, zdouble Eps, zdouble Mu,
                        double sign_field, double Iout[4])
{
  int np;
  const double *TCR = o_tcr(20, &np);
  double A[3], B[3], C, J;
  for (int i = 0; i < 3; i++) { A[i] = V2[i] - V1[i]; B[i] = V3[i] - V1[i]; }
  C = A[1]*B[2] - A[2]*B[1]; J = C*C;
  C = A[2]*B[0] - A[0]*B[2]; J += C*C;
  C = A[0]*B[1] - A[1]*B[0]; J += C*C;
  J = sqrt(J);
  memset(Iout, 0, 4*sizeof(double));
  for (int q = 0; q < np; q++) {
    double u = TCR[3*q], v = TCR[3*q+1], w = J*TCR[3*q+2];
    double X[3], f[3];
    for (int i = 0; i < 3; i++) X[i] = V1[i] + u*A[i] + v*B[i];
    for (int i = 0; i < 3; i++) f[i] = (X[i] - Q[i])*PreFac;
    zdouble EH[6];
    o_planewave_fields(E0, nHat, Omega, Eps, Mu, X, EH);
    for (int i = 0; i < 6; i++) EH[i] *= sign_field;
    zdouble e = f[0]*EH[0] + f[1]*EH[1] + f[2]*EH[2];
    zdouble h = f[0]*EH[3] + f[1]*EH[4] + f[2]*EH[5];
    Iout[0] += w*creal(e); Iout[1] += w*cimag(e);
    Iout[2] += w*creal(h); Iout[3] += w*cimag(h);
  }
}

int o_assemble_rhs_pw(OGeom *G, zdouble Omega, const zdouble E0[3],
                      const double nHat[3], zdouble *RHS)
{
  o_update_epsmu(G, Omega);
  memset(RHS, 0, sizeof(zdouble)*G->TotalBFs);
  zdouble Eps = G->EpsTF[0], Mu = G->MuTF[0];
  for (int ns = 0; ns < G->NumSurfaces; ns++) {
    OSurface *S = G->Surfaces[ns];
    int Offset = G->BFIndexOffset[ns];
    double sign;
    if (S->RegionIndices[0] == 0) sign = -1.0;
    else if (S->RegionIndices[1] == 0) sign = +1.0;
    else continue;
    for (int ne = 0; ne < S->NumEdges; ne++) {
      OEdge *E = S->Edges[ne];
      double *QP = S->Vertices + 3*E->iQP, *V1 = S->Vertices + 3*E->iV1, *V2 = S->Vertices + 3*E->iV2;
      double IP[4], IM[4] = {0, 0, 0, 0};
      panel_inner(QP, V1, V2, QP, E->Length/(2.0*S->Panels[E->iPPanel].Area), E0, nHat, Omega, Eps, Mu, sign, IP);
      if (E->iQM != -1) {
        double *QM = S->Vertices + 3*E->iQM;
        panel_inner(V1, V2, QM, QM, E->Length/(2.0*S->Panels[E->iMPanel].Area), E0, nHat, Omega, Eps, Mu, sign, IM);
      }
      zdouble EProd = (IP[0] - IM[0]) + I*(IP[1] - IM[1]);
      zdouble HProd = (IP[2] - IM[2]) + I*(IP[3] - IM[3]);
      if (S->IsPEC) RHS[Offset + ne] = EProd/O_ZVAC;
      else { RHS[Offset + 2*ne] = EProd/O_ZVAC; RHS[Offset + 2*ne + 1] = HProd; }
    }
  }
  return 0;
}

/* crossed overlap O^x_ab = \int f_a . (nHat x f_b) over one shared panel */
static double cross_overlap(OSurface *S, int np, int iQa, int iQb, double Sign, double LL)
{
  OPanel *P = S->Panels + np;
  double *Qa = S->Vertices + 3*P->VI[iQa], *Q1 = S->Vertices + 3*P->VI[(iQa+1)%3];
  double *Q2 = S->Vertices + 3*P->VI[(iQa+2)%3], *Qb = S->Vertices + 3*P->VI[iQb];
  double L1[3], L2[3], DQ[3], ZxDQ[3];
  for (int i = 0; i < 3; i++) { L1[i] = Q1[i] - Qa[i]; L2[i] = Q2[i] - Q1[i]; DQ[i] = Qa[i] - Qb[i]; }
  const double *Z = P->ZHat;
  ZxDQ[0] = Z[1]*DQ[2] - Z[2]*DQ[1]; ZxDQ[1] = Z[2]*DQ[0] - Z[0]*DQ[2]; ZxDQ[2] = Z[0]*DQ[1] - Z[1]*DQ[0];
  double tf = ((2.0*L1[0] + L2[0])*ZxDQ[0] + (2.0*L1[1] + L2[1])*ZxDQ[1] + (2.0*L1[2] + L2[2])*ZxDQ[2])/6.0;
  return Sign*LL/(2.0*P->Area)*tf;
}

void o_power(OGeom *G, zdouble Omega, const zdouble *KN, const zdouble *RHS,
             double *PAbs, double *PExt)
{
  (void)Omega;
  double pabs = 0.0, pext = 0.0;
  for (int ns = 0; ns < G->NumSurfaces; ns++) {
    OSurface *S = G->Surfaces[ns];
    int Off = G->BFIndexOffset[ns];
    for (int a = 0; a < S->NumEdges; a++) {
      if (S->IsPEC) {
        pext += 0.5*creal(-O_ZVAC*conj(KN[Off + a])*RHS[Off + a]);
      } else {
        pext += 0.5*creal(-O_ZVAC*conj(KN[Off + 2*a])*RHS[Off + 2*a]
                          + O_ZVAC*conj(KN[Off + 2*a + 1])*RHS[Off + 2*a + 1]);
      }
    }
    if (S->IsPEC) continue;
    /* absorbed power: pairs of edges sharing a panel */
    for (int a = 0; a < S->NumEdges; a++) {
      OEdge *Ea = S->Edges[a];
      int pa[2] = {Ea->iPPanel, Ea->iMPanel}, qa[2] = {Ea->PIndex, Ea->MIndex};
      for (int sa = 0; sa < 2; sa++) {
        OPanel *P = S->Panels + pa[sa];
        for (int le = 0; le < 3; le++) {
          int b = P->EI[le];
          if (b < 0) continue;
          OEdge *Eb = S->Edges[b];
          int sb = (Eb->iPPanel == pa[sa]) ? 0 : 1;
          int qb = sb == 0 ? Eb->PIndex : Eb->MIndex;
          double sign = (sa == sb) ? 1.0 : -1.0;
          double ox = cross_overlap(S, pa[sa], qa[sa], qb, sign, Ea->Length*Eb->Length);
          zdouble ka = KN[Off + 2*a], kb = KN[Off + 2*b];
          zdouble na = -O_ZVAC*KN[Off + 2*a + 1], nb = -O_ZVAC*KN[Off + 2*b + 1];
          pabs += 0.25*creal((conj(ka)*nb - conj(na)*kb)*ox);
        }
      }
    }
  }
  if (PAbs) *PAbs = pabs;
  if (PExt) *PExt = pext;
}

/* EMT scattered power (EMTPFT.cc:629-803 with the integrals of the default
   method EHDERIVATIVES2, :334-362) and absorbed power = extinction -
   scattered (:818-830).  Every surface lives in the exterior region 0
   (Sign = +1 for every pair, :640-650).  The integrals come from the
   exterior-region matrix Mext: for edges a, b with basis functions
   K (and N on a dielectric surface),
     Mext(Ka,Kb) = i w mu G_ab, Mext(Na,Nb) = -i w eps G_ab, Mext(Ka,Nb) = -i k C_ab,
   so QKK = mu w G = -i Mext(Ka,Kb), QNN = eps w G = i Mext(Na,Nb) and
   QKNmNK = i k C = -Mext(Ka,Nb).  KN bilinears (GetPFT.cc:95-132):
   k = KN(K), n = -ZVAC KN(N), and only KK when either surface is PEC.
   Self pairs use the symmetry-reduced form (:695-705); a pair of distinct
   surfaces adds the same value with both orders summed, so every ordered
   pair contributes 0.5 (Re u0KK Im QKK + Re e0NN Im QNN + Im KNmNK Im QKNmNK). */
int o_power_emt(OGeom *G, zdouble Omega, const zdouble *KN, const zdouble *RHS,
                const zdouble *Mext, int ldm, double *PAbs, double *PScat)
{
  for (int ns = 0; ns < G->NumSurfaces; ns++)
    if (G->Surfaces[ns]->RegionIndices[0] != 0) return -1;
  double pscat = 0.0;
  for (int nsa = 0; nsa < G->NumSurfaces; nsa++)
    for (int nsb = 0; nsb < G->NumSurfaces; nsb++) {
      OSurface *SA = G->Surfaces[nsa], *SB = G->Surfaces[nsb];
      int OA = G->BFIndexOffset[nsa], OB = G->BFIndexOffset[nsb];
      int both = !SA->IsPEC && !SB->IsPEC;
      for (int a = 0; a < SA->NumEdges; a++)
        for (int b = 0; b < SB->NumEdges; b++) {
          int Ka = SA->IsPEC ? OA + a : OA + 2*a, Kb = SB->IsPEC ? OB + b : OB + 2*b;
          zdouble ka = KN[Ka], kb = KN[Kb];
          zdouble u0KK = conj(ka)*kb*O_ZVAC;
          zdouble QKK = -I*Mext[(size_t)Ka + (size_t)Kb*ldm];
          double d = creal(u0KK)*cimag(QKK);
          if (both) {
            zdouble na = -O_ZVAC*KN[Ka + 1], nb = -O_ZVAC*KN[Kb + 1];
            zdouble e0NN = conj(na)*nb/O_ZVAC, KNmNK = conj(ka)*nb - conj(na)*kb;
            zdouble QNN = I*Mext[(size_t)(Ka + 1) + (size_t)(Kb + 1)*ldm];
            zdouble QKNmNK = -Mext[(size_t)Ka + (size_t)(Kb + 1)*ldm];
            d += creal(e0NN)*cimag(QNN) + cimag(KNmNK)*cimag(QKNmNK);
          }
          pscat += 0.5*d;
        }
    }
  double pabs = 0.0, pext = 0.0;
  o_power(G, Omega, KN, RHS, &pabs, &pext);
  if (PScat) *PScat = pscat;
  if (PAbs) *PAbs = pext - pscat;
  return 0;
}

/* Mie efficiencies; m_rel with creal(m_rel) < 0 selects the PEC sphere */
void o_mie_sphere(double x, zdouble m, int nmax, double *Qsca, double *Qext)
{
  int pec = creal(m) < 0;
  int nstop = (int)(x + 4.0*cbrt(x) + 2.0);
  if (nmax > 0) nstop = nmax;
  zdouble mx = m*x;
  int nmx = (int)fmax((double)nstop, cabs(mx)) + 16;
  zdouble *D = (zdouble *)calloc(nmx + 1, sizeof(zdouble));
  if (!pec) {
    D[nmx] = 0.0;
    for (int n = nmx; n >= 1; n--) D[n-1] = n/mx - 1.0/(D[n] + n/mx);
  }
  double psi0 = cos(x), psi1 = sin(x), chi0 = -sin(x), chi1 = cos(x);
  zdouble xi1 = psi1 - I*chi1;
  double qsca = 0, qext = 0;
  for (int n = 1; n <= nstop; n++) {
    double psi = (2.0*n - 1.0)*psi1/x - psi0;
    double chi = (2.0*n - 1.0)*chi1/x - chi0;
    zdouble xi = psi - I*chi;
    zdouble an, bn;
    if (pec) {
      an = (n/x*psi - psi1)/(n/x*xi - xi1);
      bn = psi/xi;
    } else {
      zdouble da = D[n]/m + n/x, db = m*D[n] + n/x;
      an = (da*psi - psi1)/(da*xi - xi1);
      bn = (db*psi - psi1)/(db*xi - xi1);
    }
    qsca += (2.0*n + 1.0)*(creal(an)*creal(an) + cimag(an)*cimag(an) + creal(bn)*creal(bn) + cimag(bn)*cimag(bn));
    qext += (2.0*n + 1.0)*creal(an + bn);
    psi0 = psi1; psi1 = psi; chi0 = chi1; chi1 = chi; xi1 = psi1 - I*chi1;
  }
  free(D);
  if (Qsca) *Qsca = 2.0/(x*x)*qsca;
  if (Qext) *Qext = 2.0/(x*x)*qext;
}
