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

void o_get_eeis(OSurface *Sa, int nea, OSurface *Sb, int neb, zdouble k,
                OFIPPICache *FC, zdouble GC[2])
{
  if (creal(k) == 0.0 && cimag(k) == 0.0) { GC[0] = GC[1] = 0.0; return; }
  OEdge *Ea = Sa->Edges[nea], *Eb = Sb->Edges[neb];
  zdouble HPP[2] = {0, 0}, HPM[2] = {0, 0}, HMP[2] = {0, 0}, HMM[2] = {0, 0};
  OPPIArgs A;
  o_ppi_init(&A);
  A.Sa = Sa; A.Sb = Sb; A.k = k; A.FC = FC;
  A.npa = Ea->iPPanel; A.iQa = Ea->PIndex; A.npb = Eb->iPPanel; A.iQb = Eb->PIndex;
  o_get_ppis(&A); HPP[0] = A.H[0]; HPP[1] = A.H[1];
  if (Eb->iMPanel != -1) {
    A.npa = Ea->iPPanel; A.iQa = Ea->PIndex; A.npb = Eb->iMPanel; A.iQb = Eb->MIndex;
    o_get_ppis(&A); HPM[0] = A.H[0]; HPM[1] = A.H[1];
  }
  if (Ea->iMPanel != -1) {
    A.npa = Ea->iMPanel; A.iQa = Ea->MIndex; A.npb = Eb->iPPanel; A.iQb = Eb->PIndex;
    o_get_ppis(&A); HMP[0] = A.H[0]; HMP[1] = A.H[1];
  }
  if (Ea->iMPanel != -1 && Eb->iMPanel != -1) {
    A.npa = Ea->iMPanel; A.iQa = Ea->MIndex; A.npb = Eb->iMPanel; A.iQb = Eb->MIndex;
    o_get_ppis(&A); HMM[0] = A.H[0]; HMM[1] = A.H[1];
  }
  double GPreFac = Ea->Length*Eb->Length;
  zdouble CPreFac = Ea->Length*Eb->Length/(I*k);
  GC[0] = GPreFac*(HPP[0] - HPM[0] - HMP[0] + HMM[0]);
  GC[1] = CPreFac*(HPP[1] - HPM[1] - HMP[1] + HMM[1]);
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

#define MADD(r, c, v) (M[(size_t)(r) + (size_t)(c)*(size_t)ldm] += (v))

/* one surface-pair block; rows restricted to edges [e0,e1) of Sa */
static void ssi_block(OGeom *G, int nsa, int nsb, zdouble Omega, OFIPPICache *FC,
                      zdouble *M, int ldm, int e0, int e1, int zero, OAsmStats *st)
{
  OSurface *Sa = G->Surfaces[nsa], *Sb = G->Surfaces[nsb];
  int RowOffset = G->BFIndexOffset[nsa], ColOffset = G->BFIndexOffset[nsb];
  int Symmetric = (nsa == nsb);
  if (zero)
    for (int c = 0; c < Sb->NumBFs; c++)
      for (int r = 0; r < Sa->NumBFs; r++) M[(size_t)(RowOffset + r) + (size_t)(ColOffset + c)*ldm] = 0.0;
  int CR[2]; double Signs[2];
  int ncr = count_common_regions(Sa, Sb, CR, Signs);
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
    for (int nsp = ns; nsp < G->NumSurfaces; nsp++)
      ssi_block(G, ns, nsp, Omega, FC, M, ldm, 0, G->Surfaces[ns]->NumEdges, 1, st);
  int N = G->TotalBFs;
  for (int nr = 1; nr < N; nr++)
    for (int nc = 0; nc < nr; nc++)
      M[(size_t)nr + (size_t)nc*ldm] = M[(size_t)nc + (size_t)nr*ldm];
  return 0;
}

int o_assemble_bem_rows(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M, int ldm, int e0, int e1)
{
  o_update_epsmu(G, Omega);
  ssi_block(G, 0, 0, Omega, FC, M, ldm, e0, e1, 0, NULL);
  return 0;
}
