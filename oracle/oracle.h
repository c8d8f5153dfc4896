/*
 * oracle.h -- CPU restatement of SCUFF-EM's BEM system path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product
 * (scuff-em_amd/) includes, links or calls it.
 *
 * Every routine restates one reference routine (file:line cited at the
 * definition); paths are relative to the reference tree root.
 */
#ifndef SCUFFX_ORACLE_H
#define SCUFFX_ORACLE_H

#include <complex.h>
#include <stddef.h>

typedef double complex zdouble;

/* ---------------------------------------------------------------------
 * geometry (libs/libscuff/libscuff.h:100-146, RWGSurface.cc:55-97,
 * InitEdgeList.cc:47-371, ParseMeshFiles.cc:69-259)
 * ------------------------------------------------------------------- */
typedef struct {
  int VI[3];
  int EI[3];
  double Centroid[3];
  double ZHat[3];
  double Radius;
  double Area;
  int Index;
  int ZHatFlipped;        /* normal reversed to point into the exterior region */
} OPanel;

typedef struct OEdge_s {
  int iV1, iV2, iQP, iQM;
  double Centroid[3];
  double Length, Radius;
  int iPPanel, iMPanel, PIndex, MIndex;
  int Index;
  struct OEdge_s *Next;
} OEdge;

typedef struct {
  int NumVertices, NumPanels, NumEdges, NumExteriorEdges;
  double *Vertices;       /* 3*NumVertices */
  OPanel *Panels;         /* NumPanels */
  OEdge **Edges;          /* NumEdges interior edges, Edges[i]->Index==i */
  OEdge **AllEdges;       /* every edge allocated (for freeing) */
  int NumAllEdges;
  int IsPEC;
  int RegionIndices[2];
  int NumBFs;
} OSurface;

/* material model for a region: eps(omega), mu(omega) */
#define OMAT_VACUUM 0
#define OMAT_CONST  1
#define OMAT_DRUDE  2   /* eps = 1 - wp^2/(w(w + i gamma)) */
#define OMAT_SIGMA  3   /* eps = Re(Eps) + i sigma/w (CONST_EPS_x_SIGMA_y, libMatProp.cc:273-280) */
typedef struct {
  int Kind;
  zdouble Eps, Mu;        /* OMAT_CONST; Re(Eps) for OMAT_SIGMA */
  double wp, gamma;       /* OMAT_DRUDE */
  double sigma;           /* OMAT_SIGMA */
} OMaterial;

#define OMAXSURF 8
#define OMAXREG  (OMAXSURF+1)
typedef struct {
  int NumSurfaces;
  OSurface *Surfaces[OMAXSURF];
  int NumRegions;
  OMaterial Materials[OMAXREG];
  int BFIndexOffset[OMAXSURF];
  int TotalBFs;
  zdouble EpsTF[OMAXREG], MuTF[OMAXREG];
  zdouble StoredOmega;
  int Mate[OMAXSURF];     /* RWGGeometry::Mate (RWGGeometry.cc:733-770), set by the caller; -1 = none */
  int RegionFilter;       /* 0: every common region; r + 1: contributions of region r only */
} OGeom;

/* mesh construction */
int  o_icosphere(int level, double radius, const double *center,
                 double **V, int *nV, int **T, int *nT);
int  o_write_gmsh(const char *fn, const double *V, int nV, const int *T, int nT);
char *o_parse_gmsh(const char *fn, double **V, int *nV, int **T, int *nT);
OSurface *o_surface_new(const double *V, int nV, const int *T, int nT,
                        int IsPEC, char **err);
void o_surface_free(OSurface *S);
OGeom *o_geom_new(void);
int  o_geom_add_surface(OGeom *G, OSurface *S, OMaterial interior);
void o_geom_orient_normals(OGeom *G);   /* RWGGeometry.cc:653-690 */
void o_geom_free(OGeom *G);
void o_update_epsmu(OGeom *G, zdouble Omega);

/* ---------------------------------------------------------------------
 * quadrature (libs/libTriInt/TCR.cc, libs/libSGJC/hcubature.c,
 * libs/libSGJC/pcubature.c)
 * ------------------------------------------------------------------- */
const double *o_tcr(int order, int *npts);
typedef void (*o_integrand)(unsigned ndim, const double *x, void *params,
                            unsigned fdim, double *fval);
int o_hcubature(unsigned fdim, o_integrand f, void *fdata, unsigned dim,
                const double *xmin, const double *xmax, size_t maxEval,
                double reqAbsError, double reqRelError,
                double *val, double *err, size_t *numEvalOut);
int o_pcubature(unsigned fdim, o_integrand f, void *fdata, unsigned dim,
                const double *xmin, const double *xmax, size_t maxEval,
                double reqAbsError, double reqRelError,
                double *val, double *err);
void o_clencurt_init(void);
extern double o_clencurt_x[];
extern double o_clencurt_w[];
#define O_CLENCURT_M 19

/* ---------------------------------------------------------------------
 * singular integrals
 * ------------------------------------------------------------------- */
typedef struct {
  double xMxpRM3[3], xXxpRM3[3];
  double uvupvpRM1[9];
  double uvupvpR1[9];
  double uvupvpR2[9];
  double x0[3];
} OQIFIPPIData;

typedef struct {
  double hTimesRM3;
  double hDotRM1, hNablaRM1, hTimesRM1;
  double hDotR0,  hNablaR0,  hTimesR0;
  double hDotR1,  hNablaR1,  hTimesR1;
  double hDotR2,  hNablaR2;
} OQDFIPPIData;

void o_compute_qifippi(double **Va, double **Vb, int ncv, OQIFIPPIData *Q,
                       size_t *nevals);

/* Taylor-Duffy (libs/libscuff/TaylorDuffy.h:34-96) */
#define OTD_RP                   0
#define OTD_HELMHOLTZ            1
#define OTD_GRADHELMHOLTZ        2
#define OTD_HIGHK_HELMHOLTZ      3
#define OTD_HIGHK_GRADHELMHOLTZ  4
#define OTD_UNITY               0
#define OTD_PMCHWG1             2
#define OTD_PMCHWC              3
#define OTD_COMMONVERTEX        1
#define OTD_COMMONEDGE          2
#define OTD_COMMONTRIANGLE      3
typedef struct {
  int WhichCase;
  int NumPKs;
  const int *PIndex;
  const int *KIndex;
  const zdouble *KParam;
  double *V1, *V2, *V3, *V2P, *V3P;
  double *Q, *QP;
  zdouble *Result, *Error;
  double AbsTol, RelTol;
  int MaxEval;
  int nCalls;
} OTDArgs;
void o_taylor_duffy_init(OTDArgs *A);
void o_taylor_duffy(OTDArgs *A);

/* FIPPI cache (libs/libscuff/FIPPICache.cc:48-187) */
typedef struct OFIPPICache OFIPPICache;
OFIPPICache *o_fippi_cache_new(void);
void o_fippi_cache_free(OFIPPICache *C);
OQIFIPPIData *o_fippi_cache_get(OFIPPICache *C, double **OVa, double **OVb,
                                int ncv);
size_t o_fippi_cache_size(OFIPPICache *C);
size_t o_fippi_cache_hits(OFIPPICache *C);

/* ---------------------------------------------------------------------
 * panel-panel interactions (libs/libscuff/PanelPanelInteractions.cc)
 * ------------------------------------------------------------------- */
#define OPPI_LOCUBATURE 0
#define OPPI_HOCUBATURE 1
#define OPPI_TD         2
#define OPPI_HKTD       3
#define OPPI_DESING     4
typedef struct {
  OSurface *Sa, *Sb;
  int npa, npb, iQa, iQb;
  zdouble k;
  int NumGradientComponents;
  OFIPPICache *FC;        /* NULL => compute FIPPIs on the fly */
  int ForceTaylorDuffy;
  int WhichAlgorithm;
  zdouble H[2];
  zdouble GradH[6];
  int NumTorqueAxes;          /* 0..3 */
  const double *GammaMatrix;  /* 9 per axis, Gamma[i + 3 j] */
  zdouble dHdT[6];
} OPPIArgs;
void o_ppi_init(OPPIArgs *A);
void o_get_ppis(OPPIArgs *A);

int  o_assess_panel_pair(OSurface *Sa, int npa, OSurface *Sb, int npb,
                         double *rRel, double **Va, double **Vb);
int  o_canonically_order(double **Va, double **Vb, int ncv,
                         double **OVa, double **OVb);

/* ---------------------------------------------------------------------
 * BEM matrix (AssembleBEMMatrix.cc:684-773,
 * SurfaceSurfaceInteractions.cc:62-334,426-585,
 * EdgeEdgeInteractions.cc:50-217)
 * M is column-major complex, M[r + c*ldm].
 * ------------------------------------------------------------------- */
typedef struct {
  unsigned long counts[5];   /* PPI algorithm histogram */
} OAsmStats;
int o_assemble_bem(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M,
                   int ldm, OAsmStats *stats);
/* one edge-edge interaction (for tests) */
void o_get_eeis(OSurface *Sa, int nea, OSurface *Sb, int neb, zdouble k,
                OFIPPICache *FC, zdouble GC[2]);
void o_get_eeis_grad(OSurface *Sa, int nea, OSurface *Sb, int neb, zdouble k,
                     OFIPPICache *FC, zdouble GC[2], zdouble *GradGC,
                     int ntorque, const double *Gamma, zdouble *dGCdT);
int o_assemble_block_grad(OGeom *G, zdouble Omega, OFIPPICache *FC, int nsa, int nsb,
                          zdouble *B, int ldb, zdouble *GradB,
                          int ntorque, const double *Gamma, zdouble *dBdT);
/* rows [r0,r1) of the upper triangle only (bounded CPU timing sample) */
int o_assemble_bem_rows(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M,
                        int ldm, int e0, int e1);
int o_assemble_bem_rows_compact(OGeom *G, zdouble Omega, OFIPPICache *FC, zdouble *M,
                                int ldm, int e0, int e1);

/* ---------------------------------------------------------------------
 * LU / solve (LAPACK zgetrf/zgetrs semantics; LBWrappers.cc:151-245)
 * ------------------------------------------------------------------- */
int o_zgetrf(int n, zdouble *A, int lda, int *ipiv);
int o_zgetrs(char trans, int n, int nrhs, const zdouble *A, int lda,
             const int *ipiv, zdouble *B, int ldb);

/* ---------------------------------------------------------------------
 * RHS, power, Mie (AssembleRHSVector.cc:190-427, OPFT.cc)
 * ------------------------------------------------------------------- */
void o_planewave_fields(const zdouble E0[3], const double nHat[3], zdouble Omega,
                        zdouble Eps, zdouble Mu, const double X[3], zdouble EH[6]);
int  o_assemble_rhs_pw(OGeom *G, zdouble Omega, const zdouble E0[3],
                       const double nHat[3], zdouble *RHS);
void o_power(OGeom *G, zdouble Omega, const zdouble *KN, const zdouble *RHS,
             double *PAbs, double *PExt);
/* energy/momentum-transfer scattered and absorbed power (EMTPFT.cc:468-803,
   default integral method EHDERIVATIVES2 :334-362) for surfaces that all
   live in the exterior region 0: Mext = the matrix assembled with the
   exterior region's contributions only */
int o_power_emt(OGeom *G, zdouble Omega, const zdouble *KN, const zdouble *RHS,
                const zdouble *Mext, int ldm, double *PAbs, double *PScat);
void o_mie_sphere(double x_size, zdouble m_rel, int nmax,
                  double *Qsca, double *Qext);

#define O_ZVAC 376.73031346177

#endif
