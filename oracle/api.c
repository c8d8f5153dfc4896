/*
 * api.c -- flat C entry points of the oracle for ctypes (tests/,
 * smoke() and the cpu_baseline leg of bench.py).  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

typedef struct { OGeom *G; OFIPPICache *FC; } oh;

void *o_api_geom_new(void)
{
  oh *h = (oh *)calloc(1, sizeof(oh));
  h->G = o_geom_new();
  h->FC = o_fippi_cache_new();
  return h;
}

void o_api_geom_free(void *p)
{
  oh *h = (oh *)p;
  if (!h) return;
  o_geom_free(h->G);
  o_fippi_cache_free(h->FC);
  free(h);
}

/* kind: 0 PEC, 1 constant eps, 2 Drude */
int o_api_add_surface(void *p, const double *V, int nV, const int *T, int nT,
                      int kind, double er, double ei, double wp, double gamma)
{
  oh *h = (oh *)p;
  char *err = NULL;
  OSurface *S = o_surface_new(V, nV, T, nT, kind == 0, &err);
  if (!S) { if (err) { fprintf(stderr, "oracle: %s\n", err); free(err); } return -1; }
  OMaterial m;
  memset(&m, 0, sizeof(m));
  /* kind 3: CONST_EPS_er_SIGMA_wp (the conductivity travels in wp) */
  m.Kind = kind == 2 ? OMAT_DRUDE : kind == 3 ? OMAT_SIGMA : OMAT_CONST;
  m.Eps = er + I*ei; m.Mu = 1.0; m.wp = wp; m.gamma = gamma;
  if (kind == 3) { m.sigma = wp; m.wp = 0.0; }
  return o_geom_add_surface(h->G, S, m);
}

int o_api_add_surface_msh(void *p, const char *fn, int kind, double er, double ei,
                          double wp, double gamma)
{
  double *V; int nV, *T, nT;
  char *err = o_parse_gmsh(fn, &V, &nV, &T, &nT);
  if (err) { fprintf(stderr, "oracle: %s\n", err); free(err); return -1; }
  int r = o_api_add_surface(p, V, nV, T, nT, kind, er, ei, wp, gamma);
  free(V); free(T);
  return r;
}

int o_api_total_bfs(void *p) { return ((oh *)p)->G->TotalBFs; }
int o_api_num_edges(void *p, int ns) { return ((oh *)p)->G->Surfaces[ns]->NumEdges; }
int o_api_num_panels(void *p, int ns) { return ((oh *)p)->G->Surfaces[ns]->NumPanels; }
int o_api_num_vertices(void *p, int ns) { return ((oh *)p)->G->Surfaces[ns]->NumVertices; }

/* edges as int[9*NE]: iV1 iV2 iQP iQM iPPanel iMPanel PIndex MIndex Index */
void o_api_edges(void *p, int ns, int *out, double *lengths)
{
  OSurface *S = ((oh *)p)->G->Surfaces[ns];
  for (int e = 0; e < S->NumEdges; e++) {
    OEdge *E = S->Edges[e];
    int *o = out + 9*e;
    o[0] = E->iV1; o[1] = E->iV2; o[2] = E->iQP; o[3] = E->iQM;
    o[4] = E->iPPanel; o[5] = E->iMPanel; o[6] = E->PIndex; o[7] = E->MIndex; o[8] = E->Index;
    if (lengths) lengths[e] = E->Length;
  }
}

void o_api_panels(void *p, int ns, double *out)
{
  OSurface *S = ((oh *)p)->G->Surfaces[ns];
  for (int q = 0; q < S->NumPanels; q++) {
    OPanel *P = S->Panels + q;
    double *o = out + 8*q;
    o[0] = P->Centroid[0]; o[1] = P->Centroid[1]; o[2] = P->Centroid[2];
    o[3] = P->ZHat[0]; o[4] = P->ZHat[1]; o[5] = P->ZHat[2];
    o[6] = P->Radius; o[7] = P->Area;
  }
}

/* panel-panel integrals on surface 0: H[4] (re,im x 2), GradH[12] */
int o_api_ppi(void *p, int npa, int iQa, int npb, int iQb, double kr, double ki,
              int ngrad, int use_cache, double *H, double *GradH)
{
  oh *h = (oh *)p;
  OPPIArgs A;
  o_ppi_init(&A);
  A.Sa = A.Sb = h->G->Surfaces[0];
  A.npa = npa; A.iQa = iQa; A.npb = npb; A.iQb = iQb;
  A.k = kr + I*ki;
  A.NumGradientComponents = ngrad;
  A.FC = use_cache ? h->FC : NULL;
  o_get_ppis(&A);
  for (int i = 0; i < 2; i++) { H[2*i] = creal(A.H[i]); H[2*i+1] = cimag(A.H[i]); }
  if (GradH) for (int i = 0; i < 6; i++) { GradH[2*i] = creal(A.GradH[i]); GradH[2*i+1] = cimag(A.GradH[i]); }
  return A.WhichAlgorithm;
}

int o_api_assemble(void *p, double wr, double wi, double *M, int ldm)
{
  oh *h = (oh *)p;
  return o_assemble_bem(h->G, wr + I*wi, h->FC, (zdouble *)M, ldm, NULL);
}

int o_api_assemble_rows(void *p, double wr, double wi, double *M, int ldm, int e0, int e1)
{
  oh *h = (oh *)p;
  return o_assemble_bem_rows(h->G, wr + I*wi, h->FC, (zdouble *)M, ldm, e0, e1);
}

int o_api_assemble_rows_compact(void *p, double wr, double wi, double *M, int ldm, int e0, int e1)
{
  oh *h = (oh *)p;
  return o_assemble_bem_rows_compact(h->G, wr + I*wi, h->FC, (zdouble *)M, ldm, e0, e1);
}

/* the matrix with region r's contributions only (r < 0: all regions) */
int o_api_assemble_region(void *p, double wr, double wi, double *M, int ldm, int r)
{
  oh *h = (oh *)p;
  h->G->RegionFilter = r < 0 ? 0 : r + 1;
  int rc = o_assemble_bem(h->G, wr + I*wi, h->FC, (zdouble *)M, ldm, NULL);
  h->G->RegionFilter = 0;
  return rc;
}

int o_api_power_emt(void *p, double wr, double wi, const double *KN, const double *RHS,
                    const double *Mext, int ldm, double *out2)
{
  oh *h = (oh *)p;
  return o_power_emt(h->G, wr + I*wi, (const zdouble *)KN, (const zdouble *)RHS, (const zdouble *)Mext, ldm,
                     out2, out2 + 1);
}

size_t o_api_fippi_count(void *p) { return o_fippi_cache_size(((oh *)p)->FC); }

/* Mate[ns] = nsm: surface ns is identical to the earlier surface nsm */
int o_api_set_mate(void *p, int ns, int nsm)
{
  OGeom *G = ((oh *)p)->G;
  if (ns < 0 || ns >= G->NumSurfaces || nsm >= ns) return -1;
  G->Mate[ns] = nsm;
  return 0;
}

/* QIFIPPI record (36 doubles: xMxpRM3[3] xXxpRM3[3] RM1[9] R1[9] R2[9] x0[3]) */
void o_api_qifippi(const double *Va9, const double *Vb9, int ncv, double *out, size_t *nevals)
{
  double va[9], vb[9];
  memcpy(va, Va9, sizeof(va)); memcpy(vb, Vb9, sizeof(vb));
  double *Va[3] = {va, va+3, va+6}, *Vb[3] = {vb, vb+3, vb+6};
  OQIFIPPIData Q;
  o_compute_qifippi(Va, Vb, ncv, &Q, nevals);
  memcpy(out, &Q, sizeof(Q));
}

/* Taylor-Duffy with user kernels (known-answer tests) */
void o_api_taylor_duffy(int ncv, int npk, const int *PIndex, const int *KIndex,
                        const double *KParam, const double *V, const double *Q,
                        double *Result, int *nCalls, int maxeval)
{
  double v[15], q[6];
  memcpy(v, V, sizeof(v));
  if (Q) memcpy(q, Q, sizeof(q));
  zdouble kp[4], R[4], E[4];
  for (int i = 0; i < npk; i++) kp[i] = KParam[2*i] + I*KParam[2*i+1];
  OTDArgs T;
  o_taylor_duffy_init(&T);
  T.WhichCase = ncv; T.NumPKs = npk; T.PIndex = PIndex; T.KIndex = KIndex; T.KParam = kp;
  T.V1 = v; T.V2 = v+3; T.V3 = v+6; T.V2P = v+9; T.V3P = v+12;
  T.Q = Q ? q : NULL; T.QP = Q ? q+3 : NULL;
  T.Result = R; T.Error = E;
  if (maxeval > 0) T.MaxEval = maxeval;
  o_taylor_duffy(&T);
  for (int i = 0; i < npk; i++) { Result[2*i] = creal(R[i]); Result[2*i+1] = cimag(R[i]); }
  if (nCalls) *nCalls = T.nCalls;
}

int o_api_hcubature(unsigned fdim, void (*f)(unsigned, const double *, void *, unsigned, double *),
                    unsigned dim, const double *xmin, const double *xmax, size_t maxEval,
                    double absTol, double relTol, double *val, double *err)
{
  return o_hcubature(fdim, (o_integrand)f, NULL, dim, xmin, xmax, maxEval, absTol, relTol, val, err, NULL);
}

int o_api_pcubature(unsigned fdim, void (*f)(unsigned, const double *, void *, unsigned, double *),
                    unsigned dim, const double *xmin, const double *xmax, size_t maxEval,
                    double absTol, double relTol, double *val, double *err)
{
  return o_pcubature(fdim, (o_integrand)f, NULL, dim, xmin, xmax, maxEval, absTol, relTol, val, err);
}

int o_api_zgetrf(int n, double *A, int lda, int *ipiv) { return o_zgetrf(n, (zdouble *)A, lda, ipiv); }
int o_api_zgetrs(char trans, int n, int nrhs, const double *A, int lda, const int *ipiv, double *B, int ldb)
{ return o_zgetrs(trans, n, nrhs, (const zdouble *)A, lda, ipiv, (zdouble *)B, ldb); }

/* plane wave E0 (complex[3] as 6 doubles), propagation nHat */
int o_api_rhs(void *p, double wr, double wi, const double *E0, const double *nHat, double *RHS)
{
  zdouble e[3] = {E0[0] + I*E0[1], E0[2] + I*E0[3], E0[4] + I*E0[5]};
  return o_assemble_rhs_pw(((oh *)p)->G, wr + I*wi, e, nHat, (zdouble *)RHS);
}

void o_api_power(void *p, double wr, double wi, const double *KN, const double *RHS, double *out)
{
  o_power(((oh *)p)->G, wr + I*wi, (const zdouble *)KN, (const zdouble *)RHS, out, out+1);
}

void o_api_mie(double x, double mr, double mi, int nmax, double *out)
{
  o_mie_sphere(x, mr + I*mi, nmax, out, out+1);
}

int o_api_icosphere(int level, double radius, const double *center, double *V, int *T)
{
  double *v; int nV, *t, nT;
  int r = o_icosphere(level, radius, center, &v, &nV, &t, &nT);
  memcpy(V, v, 3*sizeof(double)*nV); memcpy(T, t, 3*sizeof(int)*nT);
  free(v); free(t);
  return r;
}

int o_api_write_gmsh(const char *fn, const double *V, int nV, const int *T, int nT)
{ return o_write_gmsh(fn, V, nV, T, nT); }

/* triangle rule of order 4 or 20 as (u, v, w) triples; returns the count */
int o_api_tcr(int order, double *out)
{
  int n = 0;
  const double *r = o_tcr(order, &n);
  if (!r) return 0;
  if (out) memcpy(out, r, 3*sizeof(double)*n);
  return n;
}

/* block of two distinct surfaces and its displacement gradients */
int o_api_assemble_block_grad(void *p, double wr, double wi, int nsa, int nsb, double *B, int ldb, double *GradB,
                              int ntorque, const double *Gamma, double *dBdT)
{
  oh *h = (oh *)p;
  return o_assemble_block_grad(h->G, wr + I*wi, h->FC, nsa, nsb, (zdouble *)B, ldb, (zdouble *)GradB,
                               ntorque, Gamma, (zdouble *)dBdT);
}

int o_api_num_bfs(void *p, int ns) { return ((oh *)p)->G->Surfaces[ns]->NumBFs; }
