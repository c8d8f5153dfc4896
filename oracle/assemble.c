/*
 * assemble.c -- oracle: edge-edge interactions and BEM matrix assembly.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates EdgeEdgeInteractions.cc:50-217 (four panel-panel terms per
 * RWG pair), SurfaceSurfaceInteractions.cc:62-90 (common regions),
 * :106-334 (prefactors and 1x1 / 2x2 stamping), :426-585 (block driver)
 * and AssembleBEMMatrix.cc:684-773 (surface-pair loop, symmetric fill).
 * The reference spreads the edge-pair loop over OpenMP tasks; every
 * entry is written by exactly one task, so the serial order used here
 * produces the same matrix.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* GetEdgeEdgeInteractions (EdgeEdgeInteractions.cc:50-217): the four
   panel-pair integrals of two RWG functions, optionally with the three
   displacement gradients (:183-198) */
void o_get_eeis_grad(OSurface *Sa, int nea, OSurface *Sb, int neb, zdouble k,
                     OFIPPICache *FC, zdouble GC[2], zdouble *GradGC,
                     int ntorque, const double *Gamma, zdouble *dGCdT)
{
  if (!dGCdT || !Gamma) ntorque = 0;
  if (creal(k) == 0.0 && cimag(k) == 0.0) {
    GC[0] = GC[1] = 0.0;
    if (GradGC) for (int m = 0; m < 6; m++) GradGC[m] = 0.0;
    for (int m = 0; m < 2*ntorque; m++) dGCdT[m] = 0.0;
    return;
  }
  OEdge *Ea = Sa->Edges[nea], *Eb = Sb->Edges[neb];
  zdouble H[4][2], GH[4][6], TH[4][6];
  memset(H, 0, sizeof(H));
  memset(GH, 0, sizeof(GH));
  memset(TH, 0, sizeof(TH));
  OPPIArgs A;
  o_ppi_init(&A);
  A.Sa = Sa; A.Sb = Sb; A.k = k; A.FC = FC;
  A.NumGradientComponents = GradGC ? 3 : 0;
  A.NumTorqueAxes = ntorque;
  A.GammaMatrix = Gamma;
  int pa[2] = {Ea->iPPanel, Ea->iMPanel}, qa[2] = {Ea->PIndex, Ea->MIndex};
  int pb[2] = {Eb->iPPanel, Eb->iMPanel}, qb[2] = {Eb->PIndex, Eb->MIndex};
  for (int s = 0; s < 2; s++)
    for (int t = 0; t < 2; t++) {
      if (pa[s] == -1 || pb[t] == -1) continue;
      A.npa = pa[s]; A.iQa = qa[s]; A.npb = pb[t]; A.iQb = qb[t];
      o_get_ppis(&A);
      H[2*s+t][0] = A.H[0]; H[2*s+t][1] = A.H[1];
      if (GradGC) memcpy(GH[2*s+t], A.GradH, sizeof(GH[0]));
      if (ntorque) memcpy(TH[2*s+t], A.dHdT, sizeof(TH[0]));
    }
  double GPreFac = Ea->Length*Eb->Length;
  zdouble CPreFac = Ea->Length*Eb->Length/(I*k);
  GC[0] = GPreFac*(H[0][0] - H[1][0] - H[2][0] + H[3][0]);
  GC[1] = CPreFac*(H[0][1] - H[1][1] - H[2][1] + H[3][1]);
  if (GradGC)
    for (int m = 0; m < 3; m++) {
      GradGC[2*m+0] = GPreFac*(GH[0][2*m+0] - GH[1][2*m+0] - GH[2][2*m+0] + GH[3][2*m+0]);
      GradGC[2*m+1] = CPreFac*(GH[0][2*m+1] - GH[1][2*m+1] - GH[2][2*m+1] + GH[3][2*m+1]);
    }
  for (int m = 0; m < ntorque; m++) {
    dGCdT[2*m+0] = GPreFac*(TH[0][2*m+0] - TH[1][2*m+0] - TH[2][2*m+0] + TH[3][2*m+0]);
    dGCdT[2*m+1] = CPreFac*(TH[0][2*m+1] - TH[1][2*m+1] - TH[2][2*m+1] + TH[3][2*m+1]);
  }
}

void o_get_eeis(OSurface *Sa, int nea, OSurface *Sb, int neb, zdouble k,
                OFIPPICache *FC, zdouble GC[2])
{
  o_get_eeis_grad(Sa, nea, Sb, neb, k, FC, GC, NULL, 0, NULL, NULL);
}

static int count_common_regions(OSurface *Sa, OSurface *Sb, int CR[2], double Signs[2])
{
  int n = 0;
  if (Sa->RegionIndices[0] == Sb->RegionIndices[0]) { CR[n] = Sa->RegionIndices[0]; Signs[n] = +1.0; n++; }
  else if (Sa->RegionIndices[0] == Sb->RegionIndices[1]) { CR[n] = Sa->RegionIndices[0]; Signs[n] = -1.0; n++; }
  if (Sa->RegionIndices[1] == Sb->RegionIndices[0]) { CR[n] = Sa->RegionIndices[1]; Signs[n] = -1.0; n++; }
  else if (!Sa->IsPEC && !Sb->IsPEC && Sa->RegionIndices[1] == Sb->RegionIndices[1]) { CR[n] = Sa->RegionIndices[1]; Signs[n] = +1.0; n++; }
  return n;
}

static zdouble csqrt2(zdouble z)
{
  zdouble r = csqrt(z);
  return cimag(r) < 0.0 ? -r : r;
}

#define MADD(r, c, v) (M[(size_t)((r) - rbase) + (size_t)(c)*(size_t)ldm] += (v))

/* one surface-pair block; rows restricted to edges [e0,e1) of Sa */
/* rows are stored from rbase on: M[(row - rbase) + col*ldm] */
static void ssi_block(OGeom *G, int nsa, int nsb, zdouble Omega, OFIPPICache *FC,
                      zdouble *M, int ldm, int rbase, int e0, int e1, int zero, OAsmStats *st)
{
  OSurface *Sa = G->Surfaces[nsa], *Sb = G->Surfaces[nsb];
  int RowOffset = G->BFIndexOffset[nsa], ColOffset = G->BFIndexOffset[nsb];
  int Symmetric = (nsa == nsb);
  if (zero)
    for (int c = 0; c < Sb->NumBFs; c++)
      for (int r = 0; r < Sa->NumBFs; r++) M[(size_t)(RowOffset + r - rbase) + (size_t)(ColOffset + c)*ldm] = 0.0;
  int CR[2]; double Signs[2];
  int ncr = count_common_regions(Sa, Sb, CR, Signs);
  if (G->RegionFilter > 0) {
    int kept = 0;
    for (int r = 0; r < ncr; r++)
      if (CR[r] == G->RegionFilter - 1) { CR[kept] = CR[r]; Signs[kept] = Signs[r]; kept++; }
    ncr = kept;
  }
  if (ncr == 0) return;
  zdouble EpsA = G->EpsTF[CR[0]], MuA = G->MuTF[CR[0]], EpsB = 0, MuB = 0;
  double SignA = Signs[0], SignB = 0;
  if (ncr == 2) { EpsB = G->EpsTF[CR[1]]; MuB = G->MuTF[CR[1]]; SignB = Signs[1]; }
  zdouble kA = csqrt2(EpsA*MuA)*Omega;
  zdouble PF1A = SignA*I*MuA*Omega, PF2A = -SignA*I*kA, PF3A = -SignA*I*EpsA*Omega;
  zdouble kB = 0, PF1B = 0, PF2B = 0, PF3B = 0;
  if (EpsB != 0.0) {
    kB = csqrt2(EpsB*MuB)*Omega;
    PF1B = SignB*I*MuB*Omega; PF2B = -SignB*I*kB; PF3B = -SignB*I*EpsB*Omega;
  }
  (void)st;
  int NEa = Sa->NumEdges, NEb = Sb->NumEdges;
  if (e1 > NEa) e1 = NEa;
  for (int nea = e0; nea < e1; nea++)
    for (int neb = Symmetric ? nea : 0; neb < NEb; neb++) {
      zdouble GC[2];
      o_get_eeis(Sa, nea, Sb, neb, kA, FC, GC);
      if (Sa->IsPEC && Sb->IsPEC) {
        MADD(RowOffset + nea, ColOffset + neb, PF1A*GC[0]);
      } else if (Sa->IsPEC && !Sb->IsPEC) {
        int X = RowOffset + nea, Y = ColOffset + 2*neb;
        MADD(X, Y, PF1A*GC[0]); MADD(X, Y+1, PF2A*GC[1]);
      } else if (!Sa->IsPEC && Sb->IsPEC) {
        int X = RowOffset + 2*nea, Y = ColOffset + neb;
        MADD(X, Y, PF1A*GC[0]); MADD(X+1, Y, PF2A*GC[1]);
      } else {
        int X = RowOffset + 2*nea, Y = ColOffset + 2*neb;
        MADD(X, Y, PF1A*GC[0]);
        MADD(X, Y+1, PF2A*GC[1]);
        if (!Symmetric || nea != neb) MADD(X+1, Y, PF2A*GC[1]);
        MADD(X+1, Y+1, PF3A*GC[0]);
      }
      if (EpsB != 0.0) {
        o_get_eeis(Sa, nea, Sb, neb, kB, FC, GC);
        int X = RowOffset + 2*nea, Y = ColOffset + 2*neb;
        MADD(X, Y, PF1B*GC[0]);
        MADD(X, Y+1, PF2B*GC[1]);
        if (!Symmetric || nea != neb) MADD(X+1, Y, PF2B*GC[1]);
        MADD(X+1, Y+1, PF3B*GC[0]);
      }
    }
}

int o_assemble_bem(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M, int ldm, OAsmStats *st)
{
  if (ldm < G->TotalBFs) return -1;
  o_update_epsmu(G, Omega);
  if (st) memset(st, 0, sizeof(*st));
  for (int ns = 0; ns < G->NumSurfaces; ns++)
    for (int nsp = ns; nsp < G->NumSurfaces; nsp++) {
      /* a surface identical to an earlier one takes a copy of that
         surface's diagonal block (AssembleBEMMatrix.cc:728-740) */
      int nsm = (ns == nsp) ? G->Mate[ns] : -1;
      if (nsm >= 0) {
        int n = G->Surfaces[ns]->NumBFs, o = G->BFIndexOffset[ns], om = G->BFIndexOffset[nsm];
        for (int c = 0; c < n; c++)
          for (int r = 0; r <= c; r++)
            M[(size_t)(o + r) + (size_t)(o + c)*ldm] = M[(size_t)(om + r) + (size_t)(om + c)*ldm];
        continue;
      }
      ssi_block(G, ns, nsp, Omega, FC, M, ldm, 0, 0, G->Surfaces[ns]->NumEdges, 1, st);
    }
  int N = G->TotalBFs;
  for (int nr = 1; nr < N; nr++)
    for (int nc = 0; nc < nr; nc++)
      M[(size_t)nr + (size_t)nc*ldm] = M[(size_t)nc + (size_t)nr*ldm];
  return 0;
}

int o_assemble_bem_rows(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M, int ldm, int e0, int e1)
{
  o_update_epsmu(G, Omega);
  ssi_block(G, 0, 0, Omega, FC, M, ldm, 0, e0, e1, 0, NULL);
  return 0;
}

/* the same rows (upper triangle) into a compact buffer: M[(r - r0) + c*ldm]
   with r0 the first basis function of edge e0 on surface 0 */
int o_assemble_bem_rows_compact(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M, int ldm, int e0, int e1)
{
  o_update_epsmu(G, Omega);
  int r0 = G->Surfaces[0]->IsPEC ? e0 : 2*e0;
  ssi_block(G, 0, 0, Omega, FC, M, ldm, r0, e0, e1, 0, NULL);
  return 0;
}

/* AssembleBEMMatrixBlock(nsa, nsb, Omega, 0, B, GradB) for two distinct
   surfaces (SurfaceSurfaceInteractions.cc:195-327 with GradB): the block
   B (NBF_a x NBF_b, column-major, ld ldb, overwritten) and its three
   displacement derivatives GradB + m*ldb*NBF_b, m = x, y, z */
int o_assemble_block_grad(OGeom *G, zdouble Omega, OFIPPICache *FC, int nsa, int nsb,
                          zdouble *B, int ldb, zdouble *GradB,
                          int ntorque, const double *Gamma, zdouble *dBdT)
{
  if (!Gamma || !dBdT) ntorque = 0;
  if (nsa == nsb || nsa < 0 || nsb < 0 || nsa >= G->NumSurfaces || nsb >= G->NumSurfaces) return -1;
  o_update_epsmu(G, Omega);
  OSurface *Sa = G->Surfaces[nsa], *Sb = G->Surfaces[nsb];
  if (ldb < Sa->NumBFs) return -1;
  size_t bsz = (size_t)ldb*Sb->NumBFs;
  memset(B, 0, sizeof(zdouble)*bsz);
  memset(GradB, 0, 3*sizeof(zdouble)*bsz);
  if (ntorque) memset(dBdT, 0, ntorque*sizeof(zdouble)*bsz);
  int CR[2]; double Signs[2];
  int ncr = count_common_regions(Sa, Sb, CR, Signs);
  for (int r = 0; r < ncr; r++) {
    zdouble Eps = G->EpsTF[CR[r]], Mu = G->MuTF[CR[r]];
    if (Eps == 0.0) continue;
    zdouble k = csqrt2(Eps*Mu)*Omega;
    zdouble PF1 = Signs[r]*I*Mu*Omega, PF2 = -Signs[r]*I*k, PF3 = -Signs[r]*I*Eps*Omega;
    for (int nea = 0; nea < Sa->NumEdges; nea++)
      for (int neb = 0; neb < Sb->NumEdges; neb++) {
        zdouble GC[2], GG[6], GT[6];
        o_get_eeis_grad(Sa, nea, Sb, neb, k, FC, GC, GG, ntorque, Gamma, GT);
        int pecA = Sa->IsPEC, pecB = Sb->IsPEC;
        int X = pecA ? nea : 2*nea, Y = pecB ? neb : 2*neb;
#define BADD(M, x, y, v) (M)[(size_t)(x) + (size_t)(y)*ldb] += (v)
        for (int m = -1; m < 3 + ntorque; m++) {
          zdouble *T = (m < 0) ? B : (m < 3) ? GradB + (size_t)m*bsz : dBdT + (size_t)(m-3)*bsz;
          zdouble g0 = (m < 0) ? GC[0] : (m < 3) ? GG[2*m] : GT[2*(m-3)];
          zdouble g1 = (m < 0) ? GC[1] : (m < 3) ? GG[2*m+1] : GT[2*(m-3)+1];
          if (pecA && pecB) BADD(T, X, Y, PF1*g0);
          else if (pecA) { BADD(T, X, Y, PF1*g0); BADD(T, X, Y+1, PF2*g1); }
          else if (pecB) { BADD(T, X, Y, PF1*g0); BADD(T, X+1, Y, PF2*g1); }
          else {
            BADD(T, X, Y, PF1*g0); BADD(T, X, Y+1, PF2*g1);
            BADD(T, X+1, Y, PF2*g1); BADD(T, X+1, Y+1, PF3*g0);
          }
        }
#undef BADD
      }
  }
  return 0;
}
