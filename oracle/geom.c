/*
 * geom.c -- oracle: meshes, panels, RWG edges, regions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double vdist(const double *a, const double *b)
{ double d = 0.0;
  for (int i = 0; i < 3; i++) d += (a[i]-b[i])*(a[i]-b[i]);
  return sqrt(d);
}

/* ------------------------------------------------------------------ */
/* icosphere test meshes.  Not part of the reference: this is the      */
/* deterministic input generator for the benchmark workloads.          */
/* level 0 = icosahedron; each level splits every triangle into four   */
/* (a,ab,ca),(ab,b,bc),(ca,bc,c),(ab,bc,ca); midpoints are projected   */
/* onto the sphere as u=(ua+ub)/|ua+ub|.                               */
/* ------------------------------------------------------------------ */
static const int ico_faces[20][3] = {
  {0,11,5},{0,5,1},{0,1,7},{0,7,10},{0,10,11},
  {1,5,9},{5,11,4},{11,10,2},{10,7,6},{7,1,8},
  {3,9,4},{3,4,2},{3,2,6},{3,6,8},{3,8,9},
  {4,9,5},{2,4,11},{6,2,10},{8,6,7},{9,8,1}};

int o_icosphere(int level, double radius, const double *center,
                double **Vout, int *nVout, int **Tout, int *nTout)
{
  const double phi = (1.0 + sqrt(5.0)) / 2.0;
  const double base[12][3] = {
    {-1, phi, 0},{1, phi, 0},{-1,-phi, 0},{1,-phi, 0},
    {0,-1, phi},{0, 1, phi},{0,-1,-phi},{0, 1,-phi},
    {phi, 0,-1},{phi, 0, 1},{-phi, 0,-1},{-phi, 0, 1}};
  int nV = 10*(1 << (2*level)) + 2, nT = 20*(1 << (2*level));
  double *U = (double *)malloc(3*sizeof(double)*nV);
  int *T = (int *)malloc(3*sizeof(int)*nT);
  int cv = 12, ct = 20;
  for (int i = 0; i < 12; i++) {
    double n = sqrt(base[i][0]*base[i][0] + base[i][1]*base[i][1] + base[i][2]*base[i][2]);
    for (int j = 0; j < 3; j++) U[3*i+j] = base[i][j] / n;
  }
  for (int t = 0; t < 20; t++) for (int j = 0; j < 3; j++) T[3*t+j] = ico_faces[t][j];

  for (int l = 0; l < level; l++) {
    /* midpoint table keyed by (min,max) vertex pair; open addressing */
    size_t cap = 1; while (cap < (size_t)(4*ct)) cap <<= 1;
    long *keys = (long *)malloc(sizeof(long)*cap);
    int *vals = (int *)malloc(sizeof(int)*cap);
    for (size_t i = 0; i < cap; i++) keys[i] = -1;
    int *T2 = (int *)malloc(3*sizeof(int)*4*ct);
    for (int t = 0; t < ct; t++) {
      int v[3] = {T[3*t], T[3*t+1], T[3*t+2]}, m[3];
      for (int e = 0; e < 3; e++) {
        int a = v[e], b = v[(e+1)%3];
        long lo = a < b ? a : b, hi = a < b ? b : a;
        long key = lo*(long)nV + hi;
        size_t h = (size_t)((unsigned long)key * 2654435761UL) & (cap-1);
        while (keys[h] != -1 && keys[h] != key) h = (h+1) & (cap-1);
        if (keys[h] == -1) {
          double s[3];
          for (int j = 0; j < 3; j++) s[j] = U[3*a+j] + U[3*b+j];
          double n = sqrt(s[0]*s[0] + s[1]*s[1] + s[2]*s[2]);
          for (int j = 0; j < 3; j++) U[3*cv+j] = s[j] / n;
          keys[h] = key; vals[h] = cv++;
        }
        m[e] = vals[h];
      }
      /* m[0]=ab, m[1]=bc, m[2]=ca */
      int *c = T2 + 12*t;
      c[0]=v[0]; c[1]=m[0]; c[2]=m[2];
      c[3]=m[0]; c[4]=v[1]; c[5]=m[1];
      c[6]=m[2]; c[7]=m[1]; c[8]=v[2];
      c[9]=m[0]; c[10]=m[1]; c[11]=m[2];
    }
    free(keys); free(vals); free(T);
    T = T2; ct *= 4;
  }
  double *V = (double *)malloc(3*sizeof(double)*nV);
  for (int i = 0; i < nV; i++)
    for (int j = 0; j < 3; j++)
      V[3*i+j] = (center ? center[j] : 0.0) + radius*U[3*i+j];
  free(U);
  *Vout = V; *nVout = nV; *Tout = T; *nTout = nT;
  return (cv == nV && ct == nT) ? 0 : -1;
}

int o_write_gmsh(const char *fn, const double *V, int nV, const int *T, int nT)
{
  FILE *f = fopen(fn, "w");
  if (!f) return -1;
  fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", nV);
  for (int i = 0; i < nV; i++)
    fprintf(f, "%d %.17g %.17g %.17g\n", i+1, V[3*i], V[3*i+1], V[3*i+2]);
  fprintf(f, "$EndNodes\n$Elements\n%d\n", nT);
  for (int t = 0; t < nT; t++)
    fprintf(f, "%d 2 2 0 0 %d %d %d\n", t+1, T[3*t]+1, T[3*t+1]+1, T[3*t+2]+1);
  fprintf(f, "$EndElements\n");
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------ */
/* GMSH reader: restates ParseGMSHFile (ParseMeshFiles.cc:69-259):     */
/* node-index remap GMSH2HR, O(V^2) redundant-vertex merge at |dV|<1e-6*/
/* (:135-147), legacy ($NOD/$ELM) and 2.x ($Nodes/$Elements) formats,  */
/* triangles (type 2) only.  Lines are read 100 bytes at a time as the */
/* reference does.                                                      */
/* ------------------------------------------------------------------ */
static char *dupmsg(const char *m) { char *s = (char *)malloc(strlen(m)+1); strcpy(s, m); return s; }

char *o_parse_gmsh(const char *fn, double **Vout, int *nVout, int **Tout, int *nTout)
{
  FILE *f = fopen(fn, "r");
  if (!f) return dupmsg("could not open mesh file");
  char Line[100];
  int legacy = -1;
  while (legacy < 0) {
    if (!fgets(Line, 100, f)) { fclose(f); return dupmsg("failed to find node start keyword"); }
    if (!strncmp(Line, "$NOD", 4)) legacy = 1;
    else if (!strncmp(Line, "$Nodes", 6)) legacy = 0;
  }
  int NumNodes;
  if (!fgets(Line, 100, f) || !(NumNodes = atoi(Line))) { fclose(f); return dupmsg("invalid number of nodes"); }
  double *V = (double *)malloc(3*sizeof(double)*NumNodes);
  int *map = (int *)malloc(sizeof(int)*2*NumNodes);
  for (int i = 0; i < 2*NumNodes; i++) map[i] = -1;
  for (int nn = 0; nn < NumNodes; nn++) {
    int idx; double x[3];
    if (!fgets(Line, 100, f)) { fclose(f); free(V); free(map); return dupmsg("too few nodes"); }
    if (sscanf(Line, "%i %le %le %le", &idx, x, x+1, x+2) != 4) { fclose(f); free(V); free(map); return dupmsg("invalid node specification"); }
    if (idx > 2*NumNodes) { fclose(f); free(V); free(map); return dupmsg("node index out of range"); }
    V[3*nn] = x[0]; V[3*nn+1] = x[1]; V[3*nn+2] = x[2];
    map[idx] = nn;
  }
  for (int i = 0; i < NumNodes; i++)
    for (int j = i+1; j < NumNodes; j++)
      if (vdist(V+3*i, V+3*j) < 1.0e-6)
        for (int g = 0; g < 2*NumNodes; g++)
          if (map[g] == j) map[g] = i;
  if (!fgets(Line, 100, f) || strncmp(Line, legacy ? "$ENDNOD" : "$EndNodes", legacy ? 7 : 9))
    { fclose(f); free(V); free(map); return dupmsg("nodes section not terminated"); }
  if (!fgets(Line, 100, f) || strncmp(Line, legacy ? "$ELM" : "$Elements", legacy ? 4 : 9))
    { fclose(f); free(V); free(map); return dupmsg("elements section not initiated"); }
  int newfmt = !legacy;
  int NumElements;
  if (!fgets(Line, 100, f) || sscanf(Line, "%i", &NumElements) != 1 || NumElements < 0) { fclose(f); free(V); free(map); return dupmsg("invalid number of elements"); }
  int *T = (int *)malloc(3*sizeof(int)*(NumElements > 0 ? NumElements : 1));
  int nT = 0;
  for (int ne = 0; ne < NumElements; ne++) {
    if (!fgets(Line, 100, f)) { fclose(f); free(V); free(map); free(T); return dupmsg("too few elements"); }
    int VI[3] = {0,0,0}, ElType = 0;
    if (!newfmt) {
      int ElNum, RegPhys, RegElem, NodeCnt;
      if (sscanf(Line, "%i %i %i %i %i %i %i %i", &ElNum, &ElType, &RegPhys, &RegElem, &NodeCnt, VI, VI+1, VI+2) < 5)
        { fclose(f); free(V); free(map); free(T); return dupmsg("invalid element specification"); }
    } else {
      int ElNum, nTags, nRead, pos, dummy;
      if (sscanf(Line, "%i %i %i%n", &ElNum, &ElType, &nTags, &nRead) < 3)
        { fclose(f); free(V); free(map); free(T); return dupmsg("invalid element specification"); }
      pos = nRead;
      for (int t = 0; t < nTags; t++) { sscanf(Line+pos, "%i%n", &dummy, &nRead); pos += nRead; }
      int nc = sscanf(Line+pos, "%i %i %i", VI, VI+1, VI+2);
      if (ElType == 2 && nc != 3) { fclose(f); free(V); free(map); free(T); return dupmsg("invalid element specification"); }
    }
    if (ElType == 2) {
      T[3*nT] = map[VI[0]]; T[3*nT+1] = map[VI[1]]; T[3*nT+2] = map[VI[2]];
      nT++;
    }
  }
  fclose(f);
  free(map);
  *Vout = V; *nVout = NumNodes; *Tout = T; *nTout = nT;
  return 0;
}

/* ------------------------------------------------------------------ */
/* InitRWGPanel (RWGSurface.cc:55-97)                                  */
/* ------------------------------------------------------------------ */
static void init_panel(OPanel *P, const double *Vt)
{
  const double *V[3] = {Vt + 3*P->VI[0], Vt + 3*P->VI[1], Vt + 3*P->VI[2]};
  for (int i = 0; i < 3; i++) P->Centroid[i] = (V[0][i] + V[1][i] + V[2][i]) / 3.0;
  P->Radius = vdist(P->Centroid, V[0]);
  P->Radius = fmax(P->Radius, vdist(P->Centroid, V[1]));
  P->Radius = fmax(P->Radius, vdist(P->Centroid, V[2]));
  double A[3], B[3];
  for (int i = 0; i < 3; i++) { A[i] = V[1][i] - V[0][i]; B[i] = V[2][i] - V[0][i]; }
  P->ZHat[0] = A[1]*B[2] - A[2]*B[1];
  P->ZHat[1] = A[2]*B[0] - A[0]*B[2];
  P->ZHat[2] = A[0]*B[1] - A[1]*B[0];
  double n = sqrt(P->ZHat[0]*P->ZHat[0] + P->ZHat[1]*P->ZHat[1] + P->ZHat[2]*P->ZHat[2]);
  double s = (P->ZHatFlipped ? -1.0 : 1.0)/n;
  for (int i = 0; i < 3; i++) P->ZHat[i] *= s;
  P->Area = 0.5*n;
}

/* ------------------------------------------------------------------ */
/* InitEdgeList (InitEdgeList.cc:47-245): edges discovered in panel    */
/* order np, local edge ne=0..2 (vertices VI[ne],VI[(ne+1)%3]); the    */
/* first visit fixes iV1<iV2, iQP=VI[(ne+2)%3], iPPanel, PIndex; the   */
/* second visit fills iQM, iMPanel, MIndex and assigns Index in visit  */
/* order.  Exterior edges get Index=-(i+1) in list-walk order.         */
/* ------------------------------------------------------------------ */
OSurface *o_surface_new(const double *Vin, int nV, const int *T, int nT,
                        int IsPEC, char **err)
{
  OSurface *S = (OSurface *)calloc(1, sizeof(OSurface));
  S->NumVertices = nV;
  S->Vertices = (double *)malloc(3*sizeof(double)*nV);
  memcpy(S->Vertices, Vin, 3*sizeof(double)*nV);
  S->NumPanels = nT;
  S->Panels = (OPanel *)calloc(nT, sizeof(OPanel));
  S->IsPEC = IsPEC;
  for (int np = 0; np < nT; np++) {
    OPanel *P = S->Panels + np;
    P->VI[0] = T[3*np]; P->VI[1] = T[3*np+1]; P->VI[2] = T[3*np+2];
    P->Index = np;
    init_panel(P, S->Vertices);
  }
  OEdge **Lists = (OEdge **)calloc(nV, sizeof(OEdge *));
  S->AllEdges = (OEdge **)malloc(sizeof(OEdge *)*3*nT);
  int NumEdges = 0, NumTotal = 0;
  for (int np = 0; np < nT; np++)
    for (int ne = 0; ne < 3; ne++) {
      OPanel *P = S->Panels + np;
      int lo = P->VI[ne], hi = P->VI[(ne+1)%3];
      if (lo > hi) { lo = P->VI[(ne+1)%3]; hi = P->VI[ne]; }
      OEdge *E;
      for (E = Lists[lo]; E; E = E->Next) if (E->iV2 == hi) break;
      if (E) {
        if (E->iMPanel != -1) {
          if (err) *err = dupmsg("invalid mesh topology: edge shared by more than two panels");
          free(Lists); o_surface_free(S); return NULL;
        }
        E->iQM = P->VI[(ne+2)%3];
        E->iMPanel = P->Index;
        E->MIndex = (ne+2)%3;
        E->Radius = fmax(E->Radius, vdist(E->Centroid, S->Vertices + 3*E->iQM));
        E->Index = NumEdges++;
      } else {
        E = (OEdge *)calloc(1, sizeof(OEdge));
        S->AllEdges[NumTotal++] = E;
        E->Next = Lists[lo];
        Lists[lo] = E;
        E->iV1 = lo; E->iV2 = hi;
        E->iQP = P->VI[(ne+2)%3];
        E->iQM = -1;
        const double *VL = S->Vertices + 3*lo, *VG = S->Vertices + 3*hi;
        for (int i = 0; i < 3; i++) E->Centroid[i] = (VL[i] + VG[i]) / 2.0;
        E->Length = vdist(VL, VG);
        E->Radius = vdist(E->Centroid, S->Vertices + 3*E->iQP);
        E->Radius = fmax(E->Radius, vdist(E->Centroid, VL));
        E->Radius = fmax(E->Radius, vdist(E->Centroid, VG));
        E->iPPanel = P->Index;
        E->PIndex = (ne+2)%3;
        E->iMPanel = -1;
        E->Index = -1;
        E->MIndex = (ne+2)%3;
      }
    }
  S->NumAllEdges = NumTotal;
  S->NumEdges = NumEdges;
  S->Edges = (OEdge **)malloc(sizeof(OEdge *)*(NumEdges > 0 ? NumEdges : 1));
  int NumExt = 0;
  for (int nv = 0; nv < nV; nv++)
    for (OEdge *E = Lists[nv]; E; E = E->Next) {
      if (E->Index == -1) { E->Index = -(NumExt+1); NumExt++; }
      else S->Edges[E->Index] = E;
    }
  S->NumExteriorEdges = NumExt;
  free(Lists);
  for (int ne = 0; ne < NumEdges; ne++) {
    OEdge *E = S->Edges[ne];
    S->Panels[E->iPPanel].EI[E->PIndex] = ne;
    if (E->iMPanel != -1) S->Panels[E->iMPanel].EI[E->MIndex] = ne;
  }
  S->NumBFs = IsPEC ? NumEdges : 2*NumEdges;
  S->RegionIndices[0] = 0;
  S->RegionIndices[1] = -1;
  return S;
}

void o_surface_free(OSurface *S)
{
  if (!S) return;
  for (int i = 0; i < S->NumAllEdges; i++) free(S->AllEdges[i]);
  free(S->AllEdges); free(S->Edges); free(S->Panels); free(S->Vertices); free(S);
}

/* ------------------------------------------------------------------ */
/* geometry: compact objects in vacuum; each non-PEC object gets its   */
/* own interior region (RWGGeometry.cc:494-651, :708-724).             */
/* ------------------------------------------------------------------ */
OGeom *o_geom_new(void)
{
  OGeom *G = (OGeom *)calloc(1, sizeof(OGeom));
  for (int i = 0; i < OMAXSURF; i++) G->Mate[i] = -1;
  G->NumRegions = 1;
  G->Materials[0].Kind = OMAT_VACUUM;
  G->StoredOmega = 0.0;
  return G;
}

int o_geom_add_surface(OGeom *G, OSurface *S, OMaterial interior)
{
  if (G->NumSurfaces >= OMAXSURF) return -1;
  int ns = G->NumSurfaces++;
  G->Surfaces[ns] = S;
  S->RegionIndices[0] = 0;
  if (S->IsPEC) S->RegionIndices[1] = -1;
  else {
    S->RegionIndices[1] = G->NumRegions;
    G->Materials[G->NumRegions++] = interior;
  }
  G->BFIndexOffset[ns] = (ns == 0) ? 0 : G->BFIndexOffset[ns-1] + G->Surfaces[ns-1]->NumBFs;
  G->TotalBFs = G->BFIndexOffset[ns] + S->NumBFs;
  G->StoredOmega = 0.0;
  o_geom_orient_normals(G);
  return ns;
}

/* number of panels of S crossed by the ray from X toward z = -infinity
   (kdtri_surface_piercings, PointInObject.cc, without the kd-tree) */
static int surface_piercings(const OSurface *S, const double X[3])
{
  int count = 0;
  for (int p = 0; p < S->NumPanels; p++) {
    const OPanel *P = S->Panels + p;
    const double *A = S->Vertices + 3*P->VI[0], *B = S->Vertices + 3*P->VI[1], *C = S->Vertices + 3*P->VI[2];
    double d = (B[0] - A[0])*(C[1] - A[1]) - (C[0] - A[0])*(B[1] - A[1]);
    if (d == 0.0) continue;
    double u = ((X[0] - A[0])*(C[1] - A[1]) - (C[0] - A[0])*(X[1] - A[1]))/d;
    double v = ((B[0] - A[0])*(X[1] - A[1]) - (X[0] - A[0])*(B[1] - A[1]))/d;
    if (u < 0.0 || v < 0.0 || u + v > 1.0) continue;
    if (A[2] + u*(B[2] - A[2]) + v*(C[2] - A[2]) < X[2]) count++;
  }
  return count;
}

/* PointInRegion (PointInObject.cc:620-657): piercings of every non-PEC
   surface bounding region r; region 0 for an even count, an interior
   region for an odd one */
static int point_in_region(const OGeom *G, int r, const double X[3])
{
  int total = 0;
  for (int s = 0; s < G->NumSurfaces; s++) {
    const OSurface *S = G->Surfaces[s];
    if (S->IsPEC) continue;
    if (S->RegionIndices[0] == r || S->RegionIndices[1] == r) total += surface_piercings(S, X);
  }
  return r == 0 ? (total % 2 == 0) : (total % 2 == 1);
}

/* panel normals into each surface's exterior region (RWGGeometry.cc:653-690):
   the point half a panel radius along the normal must lie in the exterior
   region; if it lies in the interior region the normal is flipped */
void o_geom_orient_normals(OGeom *G)
{
  for (int s = 0; s < G->NumSurfaces; s++) {
    OSurface *S = G->Surfaces[s];
    for (int p = 0; p < S->NumPanels; p++) {
      OPanel *P = S->Panels + p;
      P->ZHatFlipped = 0;
      init_panel(P, S->Vertices);
      double X[3];
      for (int i = 0; i < 3; i++) X[i] = P->Centroid[i] + 0.5*P->Radius*P->ZHat[i];
      if (point_in_region(G, S->RegionIndices[0], X)) continue;
      if (point_in_region(G, S->RegionIndices[1], X)) {
        P->ZHatFlipped = 1;
        for (int i = 0; i < 3; i++) P->ZHat[i] = -P->ZHat[i];
      }
    }
  }
}

void o_geom_free(OGeom *G)
{
  if (!G) return;
  for (int i = 0; i < G->NumSurfaces; i++) o_surface_free(G->Surfaces[i]);
  free(G);
}

/* UpdateCachedEpsMuValues (RWGGeometry.cc:1093-1112) -> GetEpsMu */
void o_update_epsmu(OGeom *G, zdouble Omega)
{
  for (int r = 0; r < G->NumRegions; r++) {
    OMaterial *m = G->Materials + r;
    if (m->Kind == OMAT_VACUUM) { G->EpsTF[r] = 1.0; G->MuTF[r] = 1.0; }
    else if (m->Kind == OMAT_CONST) { G->EpsTF[r] = m->Eps; G->MuTF[r] = m->Mu; }
    else if (m->Kind == OMAT_SIGMA) { G->EpsTF[r] = creal(m->Eps) + I*m->sigma/Omega; G->MuTF[r] = 1.0; }
    else { G->EpsTF[r] = 1.0 - m->wp*m->wp / (Omega*(Omega + I*m->gamma)); G->MuTF[r] = 1.0; }
  }
  G->StoredOmega = Omega;
}
