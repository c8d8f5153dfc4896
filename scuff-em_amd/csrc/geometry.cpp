// geometry.cpp -- host-side mesh topology (see geometry.h).
#include "geometry.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <strings.h>
#include <unordered_map>

namespace scuffx {

static double dist3(const double *a, const double *b)
{
  double d = 0.0;
  for (int i = 0; i < 3; i++) d += (a[i] - b[i]) * (a[i] - b[i]);
  return std::sqrt(d);
}

void Material::eps_mu(cdouble omega, cdouble *e, cdouble *m) const
{
  switch (kind) {
    case MAT_VACUUM: *e = 1.0; *m = 1.0; break;
    case MAT_CONST: *e = eps; *m = mu; break;
    case MAT_SIGMA: *e = eps.real() + cdouble(0.0, 1.0) * sigma / omega; *m = mu; break;
    case MAT_EXPR: {
      // user-defined expressions take w in rad/s (omega in units of c / 1 um,
      // MatProp::FreqUnit, libMatProp.cc:44)
      const cdouble w = omega * 2.99792458e14;
      *e = epsx ? eval_cexpr(*epsx, w) : cdouble(1.0);
      *m = mux ? eval_cexpr(*mux, w) : cdouble(1.0);
      break;
    }
    default: *e = 1.0 - wp * wp / (omega * (omega + cdouble(0.0, gamma))); *m = 1.0; break;
  }
}

// Panel geometry: centroid, bounding radius, normal and area
// (InitRWGPanel, RWGSurface.cc:55-97).
static void init_panel(Panel *P, const double *V)
{
  const double *X[3] = {V + 3 * P->VI[0], V + 3 * P->VI[1], V + 3 * P->VI[2]};
  for (int i = 0; i < 3; i++) P->Centroid[i] = (X[0][i] + X[1][i] + X[2][i]) / 3.0;
  P->Radius = dist3(P->Centroid, X[0]);
  P->Radius = std::fmax(P->Radius, dist3(P->Centroid, X[1]));
  P->Radius = std::fmax(P->Radius, dist3(P->Centroid, X[2]));
  double A[3], B[3];
  for (int i = 0; i < 3; i++) { A[i] = X[1][i] - X[0][i]; B[i] = X[2][i] - X[0][i]; }
  P->ZHat[0] = A[1] * B[2] - A[2] * B[1];
  P->ZHat[1] = A[2] * B[0] - A[0] * B[2];
  P->ZHat[2] = A[0] * B[1] - A[1] * B[0];
  double n = std::sqrt(P->ZHat[0] * P->ZHat[0] + P->ZHat[1] * P->ZHat[1] + P->ZHat[2] * P->ZHat[2]);
  double s = (P->ZHatFlipped ? -1.0 : 1.0) / n;
  for (int i = 0; i < 3; i++) P->ZHat[i] *= s;
  P->Area = 0.5 * n;
}

// RWG edges.  Edges are met panel by panel, local edge ne spanning
// VI[ne]..VI[(ne+1)%3].  The first panel to touch an edge is its
// positive panel (QP = opposite vertex); the second is its negative
// panel, and the edge receives the next interior index at that moment
// (InitEdgeList.cc:126-245).  Exterior edges are numbered -(i+1) in the
// order of a walk over vertices, newest edge first.
std::string build_surface(const double *Vin, int nV, const int *T, int nT,
                          bool pec, Surface *S)
{
  S->V.assign(Vin, Vin + 3 * (size_t)nV);
  S->Panels.resize(nT);
  S->IsPEC = pec;
  for (int p = 0; p < nT; p++) {
    Panel &P = S->Panels[p];
    for (int j = 0; j < 3; j++) {
      P.VI[j] = T[3 * p + j];
      P.EI[j] = 0;
      if (P.VI[j] < 0 || P.VI[j] >= nV) return "panel vertex index out of range";
    }
    init_panel(&P, S->V.data());
  }
  // per-vertex chains of edges keyed by their lesser vertex; a chain
  // is walked newest-first
  struct Tmp { Edge e; int next; };
  std::vector<Tmp> all;
  all.reserve(3 * (size_t)nT / 2 + 16);
  std::vector<int> head(nV, -1);
  int numInterior = 0;
  for (int p = 0; p < nT; p++)
    for (int ne = 0; ne < 3; ne++) {
      const Panel &P = S->Panels[p];
      int lo = P.VI[ne], hi = P.VI[(ne + 1) % 3];
      if (lo > hi) std::swap(lo, hi);
      int k = head[lo];
      while (k >= 0 && all[k].e.iV2 != hi) k = all[k].next;
      if (k >= 0) {
        Edge &E = all[k].e;
        if (E.iMPanel != -1) return "invalid mesh topology: an edge belongs to more than two panels";
        E.iQM = P.VI[(ne + 2) % 3];
        E.iMPanel = p;
        E.MIndex = (ne + 2) % 3;
        E.Radius = std::fmax(E.Radius, dist3(E.Centroid, S->V.data() + 3 * E.iQM));
        E.Index = numInterior++;
      } else {
        Tmp t;
        Edge &E = t.e;
        E.iV1 = lo; E.iV2 = hi;
        E.iQP = P.VI[(ne + 2) % 3];
        E.iQM = -1;
        const double *VL = S->V.data() + 3 * lo, *VG = S->V.data() + 3 * hi;
        for (int i = 0; i < 3; i++) E.Centroid[i] = (VL[i] + VG[i]) / 2.0;
        E.Length = dist3(VL, VG);
        E.Radius = dist3(E.Centroid, S->V.data() + 3 * E.iQP);
        E.Radius = std::fmax(E.Radius, dist3(E.Centroid, VL));
        E.Radius = std::fmax(E.Radius, dist3(E.Centroid, VG));
        E.iPPanel = p;
        E.PIndex = (ne + 2) % 3;
        E.iMPanel = -1;
        E.MIndex = (ne + 2) % 3;
        E.Index = -1;
        t.next = head[lo];
        head[lo] = (int)all.size();
        all.push_back(t);
      }
    }
  S->Edges.assign(numInterior, Edge());
  int numExt = 0;
  for (int v = 0; v < nV; v++)
    for (int k = head[v]; k >= 0; k = all[k].next) {
      Edge &E = all[k].e;
      if (E.Index == -1) E.Index = -(++numExt);
      else S->Edges[E.Index] = E;
    }
  S->NumExteriorEdges = numExt;
  for (int e = 0; e < numInterior; e++) {
    const Edge &E = S->Edges[e];
    S->Panels[E.iPPanel].EI[E.PIndex] = e;
    if (E.iMPanel != -1) S->Panels[E.iMPanel].EI[E.MIndex] = e;
  }
  S->NumBFs = pec ? numInterior : 2 * numInterior;
  return "";
}

std::vector<int> find_mates(const Geometry &G)
{
  const int NS = (int)G.Surfaces.size();
  std::vector<int> mate(NS, -1);
  for (int ns = 1; ns < NS; ns++) {
    const Surface &S = G.Surfaces[ns];
    for (int nsp = 0; nsp < ns && mate[ns] == -1; nsp++) {
      const Surface &P = G.Surfaces[nsp];
      if (S.MeshFile.empty() || S.MeshFile != P.MeshFile) continue;
      if (S.RegionIndices[0] != P.RegionIndices[0]) continue;
      if ((S.IsPEC && P.IsPEC) || (!S.IsPEC && !P.IsPEC && S.MaterialName == P.MaterialName)) mate[ns] = nsp;
    }
  }
  return mate;
}

int add_surface(Geometry *G, Surface &&S, const Material &interior)
{
  S.RegionIndices[0] = 0;
  if (S.IsPEC) S.RegionIndices[1] = -1;
  else {
    S.RegionIndices[1] = (int)G->Regions.size();
    G->Regions.push_back(interior);
  }
  int ns = (int)G->Surfaces.size();
  G->BFIndexOffset.push_back(G->TotalBFs);
  G->TotalBFs += S.NumBFs;
  G->Surfaces.push_back(std::move(S));
  orient_normals(G);
  return ns;
}

// Vertical-ray piercing counts of one surface's panels: triangles binned
// by their xy bounding boxes on a uniform grid, a query walks its cell.
namespace {
struct PierceGrid {
  const Surface *S = nullptr;
  double lo[2] = {0, 0}, h[2] = {1, 1};
  int n = 1;
  std::vector<int> start, tri;
  explicit PierceGrid(const Surface &Sf) : S(&Sf)
  {
    const int nP = Sf.NumPanels();
    if (nP == 0) return;
    double hi[2] = {-1e300, -1e300};
    lo[0] = lo[1] = 1e300;
    for (int v = 0; v < Sf.NumVertices(); v++)
      for (int j = 0; j < 2; j++) { lo[j] = std::min(lo[j], Sf.V[3 * v + j]); hi[j] = std::max(hi[j], Sf.V[3 * v + j]); }
    n = std::max(1, (int)std::sqrt((double)nP));
    for (int j = 0; j < 2; j++) h[j] = std::max(hi[j] - lo[j], 1e-300) / n;
    std::vector<std::vector<int>> cells((size_t)n * n);
    for (int p = 0; p < nP; p++) {
      int c0[2], c1[2];
      for (int j = 0; j < 2; j++) {
        double a = 1e300, b = -1e300;
        for (int k = 0; k < 3; k++) { double x = Sf.V[3 * Sf.Panels[p].VI[k] + j]; a = std::min(a, x); b = std::max(b, x); }
        c0[j] = cell(a, j); c1[j] = cell(b, j);
      }
      for (int x = c0[0]; x <= c1[0]; x++)
        for (int y = c0[1]; y <= c1[1]; y++) cells[(size_t)x * n + y].push_back(p);
    }
    start.assign((size_t)n * n + 1, 0);
    for (size_t c = 0; c < cells.size(); c++) start[c + 1] = start[c] + (int)cells[c].size();
    tri.reserve(start.back());
    for (auto &c : cells) tri.insert(tri.end(), c.begin(), c.end());
  }
  int cell(double x, int j) const { return std::max(0, std::min(n - 1, (int)std::floor((x - lo[j]) / h[j]))); }
  // triangles crossed by the ray from X toward z = -infinity
  int piercings(const double X[3]) const
  {
    if (start.empty()) return 0;
    const int c = cell(X[0], 0) * n + cell(X[1], 1);
    int count = 0;
    for (int q = start[c]; q < start[c + 1]; q++) {
      const Panel &P = S->Panels[tri[q]];
      const double *A = &S->V[3 * P.VI[0]], *B = &S->V[3 * P.VI[1]], *C = &S->V[3 * P.VI[2]];
      const double d = (B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1]);
      if (d == 0.0) continue;                         // vertical panel
      const double u = ((X[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (X[1] - A[1])) / d;
      const double v = ((B[0] - A[0]) * (X[1] - A[1]) - (X[0] - A[0]) * (B[1] - A[1])) / d;
      if (u < 0.0 || v < 0.0 || u + v > 1.0) continue;
      const double z = A[2] + u * (B[2] - A[2]) + v * (C[2] - A[2]);
      if (z < X[2]) count++;
    }
    return count;
  }
};
}  // namespace

// A point just off the panel along its normal should lie in the surface's
// exterior region; if it lies in the interior instead, the normal is
// flipped.  Region membership counts the piercings of a vertical ray with
// every non-PEC surface bounding the region: region 0 for an even count,
// an interior region for an odd one (PointInRegion, PointInObject.cc:620-657).
void orient_normals(Geometry *G)
{
  const int NS = (int)G->Surfaces.size();
  std::vector<PierceGrid> grids;
  grids.reserve(NS);
  for (int s = 0; s < NS; s++) grids.emplace_back(G->Surfaces[s]);
  auto in_region = [&](int r, const double X[3]) {
    int total = 0;
    for (int s = 0; s < NS; s++) {
      const Surface &S = G->Surfaces[s];
      if (S.IsPEC) continue;
      if (S.RegionIndices[0] == r || S.RegionIndices[1] == r) total += grids[s].piercings(X);
    }
    return r == 0 ? (total % 2 == 0) : (total % 2 == 1);
  };
  for (int s = 0; s < NS; s++) {
    Surface &S = G->Surfaces[s];
    for (Panel &P : S.Panels) {
      P.ZHatFlipped = false;
      init_panel(&P, S.V.data());
      double X[3];
      for (int i = 0; i < 3; i++) X[i] = P.Centroid[i] + 0.5 * P.Radius * P.ZHat[i];
      if (in_region(S.RegionIndices[0], X)) continue;
      if (in_region(S.RegionIndices[1], X)) {
        P.ZHatFlipped = true;
        for (double &z : P.ZHat) z = -z;
      }
    }
  }
}

// ---------------------------------------------------------------------
// GMSH reader.  Node numbers are remapped to storage order; vertices
// closer than 1e-6 are merged onto the earlier one (O(V^2), as the
// reference does); only triangles (element type 2) are kept.
// ---------------------------------------------------------------------
std::string read_gmsh(const std::string &fn, std::vector<double> *V, std::vector<int> *T)
{
  FILE *f = std::fopen(fn.c_str(), "r");
  if (!f) return "could not open " + fn;
  char Line[100];
  int legacy = -1;
  while (legacy < 0) {
    if (!std::fgets(Line, 100, f)) { std::fclose(f); return fn + ": no node section"; }
    if (!std::strncmp(Line, "$NOD", 4)) legacy = 1;
    else if (!std::strncmp(Line, "$Nodes", 6)) legacy = 0;
  }
  int NumNodes = 0;
  if (!std::fgets(Line, 100, f) || !(NumNodes = std::atoi(Line))) { std::fclose(f); return fn + ": invalid number of nodes"; }
  V->assign(3 * (size_t)NumNodes, 0.0);
  std::vector<int> map(2 * (size_t)NumNodes, -1);
  for (int nn = 0; nn < NumNodes; nn++) {
    int idx; double x[3];
    if (!std::fgets(Line, 100, f) || std::sscanf(Line, "%i %le %le %le", &idx, x, x + 1, x + 2) != 4) { std::fclose(f); return fn + ": invalid node"; }
    if (idx < 0 || idx >= 2 * NumNodes) { std::fclose(f); return fn + ": node index out of range"; }
    for (int j = 0; j < 3; j++) (*V)[3 * nn + j] = x[j];
    map[idx] = nn;
  }
  for (int i = 0; i < NumNodes; i++)
    for (int j = i + 1; j < NumNodes; j++)
      if (dist3(V->data() + 3 * i, V->data() + 3 * j) < 1.0e-6)
        for (auto &m : map) if (m == j) m = i;
  if (!std::fgets(Line, 100, f) || std::strncmp(Line, legacy ? "$ENDNOD" : "$EndNodes", legacy ? 7 : 9)) { std::fclose(f); return fn + ": nodes section not terminated"; }
  if (!std::fgets(Line, 100, f) || std::strncmp(Line, legacy ? "$ELM" : "$Elements", legacy ? 4 : 9)) { std::fclose(f); return fn + ": no elements section"; }
  int NumElements = 0;
  if (!std::fgets(Line, 100, f) || std::sscanf(Line, "%i", &NumElements) != 1 || NumElements < 0) { std::fclose(f); return fn + ": invalid number of elements"; }
  T->clear();
  for (int ne = 0; ne < NumElements; ne++) {
    if (!std::fgets(Line, 100, f)) { std::fclose(f); return fn + ": too few elements"; }
    int VI[3] = {0, 0, 0}, ElType = 0;
    if (legacy) {
      int a, b, c, d;
      if (std::sscanf(Line, "%i %i %i %i %i %i %i %i", &a, &ElType, &b, &c, &d, VI, VI + 1, VI + 2) < 5) { std::fclose(f); return fn + ": invalid element"; }
    } else {
      int ElNum, nTags, nRead, pos, dummy;
      if (std::sscanf(Line, "%i %i %i%n", &ElNum, &ElType, &nTags, &nRead) < 3) { std::fclose(f); return fn + ": invalid element"; }
      pos = nRead;
      for (int t = 0; t < nTags; t++) { std::sscanf(Line + pos, "%i%n", &dummy, &nRead); pos += nRead; }
      int nc = std::sscanf(Line + pos, "%i %i %i", VI, VI + 1, VI + 2);
      if (ElType == 2 && nc != 3) { std::fclose(f); return fn + ": invalid triangle"; }
    }
    if (ElType == 2)
      for (int j = 0; j < 3; j++) {
        if (VI[j] < 0 || VI[j] >= 2 * NumNodes || map[VI[j]] < 0) { std::fclose(f); return fn + ": triangle references an unknown node"; }
        T->push_back(map[VI[j]]);
      }
  }
  std::fclose(f);
  return "";
}

std::string parse_material(const std::string &s, Material *m, bool *pec)
{
  *pec = false;
  m->kind = MAT_CONST; m->eps = 1.0; m->mu = 1.0;
  if (s == "PEC") { *pec = true; return ""; }
  if (s == "VACUUM") { m->kind = MAT_VACUUM; return ""; }
  if (!strncasecmp(s.c_str(), "DRUDE_", 6)) {
    m->kind = MAT_DRUDE;
    if (std::sscanf(s.c_str() + 6, "%le_%le", &m->wp, &m->gamma) != 2) return "bad DRUDE material " + s;
    return "";
  }
  if (!strncasecmp(s.c_str(), "CONST_", 6)) {
    // CONST_EPS_x[_MU_y][_SIGMA_z] (libs/libMatProp/libMatProp.cc:201-290)
    std::stringstream ss(s.substr(6));
    std::string tok;
    std::vector<std::string> toks;
    while (std::getline(ss, tok, '_')) toks.push_back(tok);
    double sigma = 0.0;
    for (size_t i = 0; i < toks.size(); i++) {
      if (i + 1 >= toks.size()) return "no value in material " + s;
      double r, im; char c1, c2;
      int n = std::sscanf(toks[i + 1].c_str(), "%le%c%le%c", &r, &c1, &im, &c2);
      cdouble val;
      if (n == 1) val = r;
      else if (n == 4 && c1 == '+' && std::tolower(c2) == 'i') val = cdouble(r, im);
      else if (n == 4 && c1 == '-' && std::tolower(c2) == 'i') val = cdouble(r, -im);
      else if (!strcasecmp(toks[i].c_str(), "SIGMA") && n >= 1) val = r;
      else return "bad value in material " + s;
      if (!strcasecmp(toks[i].c_str(), "EPS")) m->eps = val;
      else if (!strcasecmp(toks[i].c_str(), "MU")) m->mu = val;
      else if (!strcasecmp(toks[i].c_str(), "SIGMA")) sigma = val.real();
      else return "invalid token in material " + s;
      i++;
    }
    if (sigma != 0.0) {
      // eps(w) = Re(eps) + i*sigma/w in units of c/1um; the reference
      // prints Re(eps) and sigma with %e before re-parsing them
      char buf[64];
      std::snprintf(buf, sizeof buf, "%e", m->eps.real());
      m->eps = std::atof(buf);
      std::snprintf(buf, sizeof buf, "%e", sigma);
      m->sigma = std::atof(buf);
      m->kind = MAT_SIGMA;
    }
    return "";
  }
  return "unsupported material " + s;
}

// ---------------------------------------------------------------------
// rigid transformations (GTransformation.cc).  An OBJECT's DISPLACED /
// ROTATED lines compose on the left into one transformation applied to
// the vertices before the panels and edges are built (RWGSurface.cc:
// 165-186, :427-435); .trans files move surfaces of a built geometry.
// ---------------------------------------------------------------------
// v <- (u.v) u + cos(t) (v - (u.v) u) + sin(t) (u x v), in the
// reference's operation order so that rotated meshes agree to the bit
static void rotate_vec(const double u[3], double c, double sn, double v[3])
{
  double ud = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
  double par[3] = {u[0] * ud, u[1] * ud, u[2] * ud};
  double perp[3] = {v[0] - par[0], v[1] - par[1], v[2] - par[2]};
  double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  for (int i = 0; i < 3; i++) v[i] = par[i] + (perp[i] * c + cr[i] * sn);
}

void GTransform::displace(const double d[3]) { for (int i = 0; i < 3; i++) DX[i] += d[i]; }

void GTransform::rotate(const double axis[3], double degrees)
{
  double n = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  double u[3] = {axis[0] / n, axis[1] / n, axis[2] / n};
  double t = degrees * (3.14159265358979323846 / 180.0);
  double c = std::cos(t), sn = std::sin(t);
  rotate_vec(u, c, sn, DX);
  for (int j = 0; j < 3; j++) {
    double col[3] = {M[0][j], M[1][j], M[2][j]};
    rotate_vec(u, c, sn, col);
    for (int i = 0; i < 3; i++) M[i][j] = col[i];
  }
}

void GTransform::compose(const GTransform &B)
{
  double d[3];
  for (int i = 0; i < 3; i++) d[i] = B.M[i][0] * DX[0] + B.M[i][1] * DX[1] + B.M[i][2] * DX[2] + B.DX[i];
  double r[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      r[i][j] = 0;
      for (int k = 0; k < 3; k++) r[i][j] += B.M[i][k] * M[k][j];
    }
  std::memcpy(DX, d, sizeof DX);
  std::memcpy(M, r, sizeof M);
}

void GTransform::apply(double *X, size_t npoints) const
{
  for (size_t n = 0; n < npoints; n++) {
    double *x = X + 3 * n;
    double x0 = x[0], x1 = x[1], x2 = x[2];
    for (int i = 0; i < 3; i++) x[i] = M[i][0] * x0 + M[i][1] * x1 + M[i][2] * x2 + DX[i];
  }
}

void GTransform::unapply(double *X, size_t npoints) const
{
  for (size_t n = 0; n < npoints; n++) {
    double *x = X + 3 * n;
    double x0 = x[0] - DX[0], x1 = x[1] - DX[1], x2 = x[2] - DX[2];
    for (int i = 0; i < 3; i++) x[i] = M[0][i] * x0 + M[1][i] * x1 + M[2][i] * x2;
  }
}

std::string parse_transform(const std::vector<std::string> &tok, size_t first, GTransform *out, size_t *consumed)
{
  char *end;
  auto num = [&](const std::string &s, double *v) {
    *v = std::strtod(s.c_str(), &end);
    return end != s.c_str() && *end == 0;
  };
  const size_t n = tok.size() - first;
  *consumed = 0;
  if (n == 0) return "no transformation specified";
  const std::string &kw = tok[first];
  if (!strcasecmp(kw.c_str(), "DISPLACED") || !strcasecmp(kw.c_str(), "DISP")) {
    double d[3];
    if (n < 4) return "too few values specified for " + kw;
    for (int i = 0; i < 3; i++)
      if (!num(tok[first + 1 + i], d + i)) return "bad value specified for " + kw;
    out->displace(d);
    *consumed = 4;
    return "";
  }
  if (!strcasecmp(kw.c_str(), "ROTATED") || !strcasecmp(kw.c_str(), "ROT")) {
    double t, a[3];
    if (n < 6) return "too few values specified for " + kw;
    if (strcasecmp(tok[first + 2].c_str(), "ABOUT")) return "invalid syntax for " + kw + " statement";
    if (!num(tok[first + 1], &t)) return "bad value specified for " + kw;
    for (int i = 0; i < 3; i++)
      if (!num(tok[first + 3 + i], a + i)) return "bad value specified for " + kw;
    out->rotate(a, t);
    *consumed = 6;
    return "";
  }
  return "unknown keyword " + kw;
}

static std::vector<std::string> tokenize(const std::string &line)
{
  std::stringstream ss(line);
  std::vector<std::string> tok;
  std::string w;
  while (ss >> w) tok.push_back(w);
  return tok;
}

std::string read_trans_file(const std::string &fn, std::vector<GTComplex> *out)
{
  out->clear();
  if (fn.empty()) {
    GTComplex C;
    C.Tag = "DEFAULT";
    out->push_back(C);
    return "";
  }
  std::ifstream in(fn);
  if (!in) return "could not open file " + fn;
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    lineno++;
    std::vector<std::string> tok = tokenize(line);
    if (tok.empty() || tok[0][0] == '#') continue;
    std::string where = fn + ":" + std::to_string(lineno) + ": ";
    GTComplex C;
    int cur = -1;   // the transformation later DISP / ROT apply to
    auto new_gt = [&](const std::string &label) {
      C.GTs.push_back(GTransform());
      C.Labels.push_back(label);
      cur = (int)C.GTs.size() - 1;
    };
    if (!strcasecmp(tok[0].c_str(), "TRANSFORMATION")) {
      C.Tag = tok.size() > 1 ? tok[1] : "DEFAULT";
      bool closed = false;
      while (!closed && std::getline(in, line)) {
        lineno++;
        where = fn + ":" + std::to_string(lineno) + ": ";
        std::vector<std::string> t = tokenize(line);
        if (t.empty() || t[0][0] == '#') continue;
        if (!strcasecmp(t[0].c_str(), "OBJECT") || !strcasecmp(t[0].c_str(), "SURFACE")) {
          if (t.size() != 2) return where + "OBJECT/SURFACE keyword requires one argument";
          new_gt(t[1]);
        } else if (!strcasecmp(t[0].c_str(), "ENDTRANSFORMATION")) {
          closed = true;
        } else {
          if (cur < 0) new_gt("");
          size_t used;
          std::string e = parse_transform(t, 0, &C.GTs[cur], &used);
          if (!e.empty()) return where + e;
          if (used != t.size()) return where + "junk at end of line";
        }
      }
      if (!closed) return fn + ": unexpected end of file in TRANSFORMATION " + C.Tag;
    } else if (!strcasecmp(tok[0].c_str(), "TRANS")) {
      if (tok.size() < 2) return where + "TRANS without a tag";
      C.Tag = tok[1];
      for (size_t nt = 2; nt < tok.size();) {
        if (!strcasecmp(tok[nt].c_str(), "OBJECT") || !strcasecmp(tok[nt].c_str(), "SURFACE")) {
          if (nt + 1 == tok.size()) return where + "empty " + tok[nt] + " statement";
          new_gt(tok[nt + 1]);
          nt += 2;
        } else if (!strcasecmp(tok[nt].c_str(), "DISP") || !strcasecmp(tok[nt].c_str(), "DISPLACED") ||
                   !strcasecmp(tok[nt].c_str(), "ROT") || !strcasecmp(tok[nt].c_str(), "ROTATED")) {
          if (cur < 0) new_gt("");
          size_t used;
          std::string e = parse_transform(tok, nt, &C.GTs[cur], &used);
          if (!e.empty()) return where + e;
          nt += used;
        } else {
          return where + "unknown keyword " + tok[nt];
        }
      }
    } else {
      return where + "unknown token " + tok[0];
    }
    out->push_back(C);
  }
  return "";
}

int surface_by_label(const Geometry &G, const std::string &label)
{
  for (size_t ns = 0; ns < G.Surfaces.size(); ns++)
    if (!strcasecmp(label.c_str(), G.Surfaces[ns].Label.c_str())) return (int)ns;
  return -1;
}

std::string check_gtc(const Geometry &G, const GTComplex &C)
{
  for (size_t k = 0; k < C.GTs.size(); k++) {
    if (C.Labels[k].empty()) return "no OBJECT/SURFACE specified for GTransformation " + C.Tag;
    if (surface_by_label(G, C.Labels[k]) < 0)
      return "unknown OBJECT/SURFACE " + C.Labels[k] + " specified for GTransformation " + C.Tag;
  }
  return "";
}

static void reinit_panels(Surface *S)
{
  for (Panel &P : S->Panels) init_panel(&P, S->V.data());
}

void transform_surface(Geometry *G, int ns, const GTransform &D)
{
  Surface &S = G->Surfaces[ns];
  D.apply(S.V.data(), S.V.size() / 3);
  for (Edge &E : S.Edges) D.apply(E.Centroid, 1);
  reinit_panels(&S);
  S.GT.compose(D);
  if ((int)G->SurfaceMoved.size() != (int)G->Surfaces.size()) G->SurfaceMoved.assign(G->Surfaces.size(), 0);
  G->SurfaceMoved[ns] = 1;
}

void transform_geometry(Geometry *G, const GTComplex &C)
{
  G->SurfaceMoved.assign(G->Surfaces.size(), 0);
  for (size_t k = 0; k < C.GTs.size(); k++) {
    const int ns = C.Labels[k].empty() ? -1 : surface_by_label(*G, C.Labels[k]);
    if (ns < 0) continue;   // the reference warns and ignores it
    transform_surface(G, ns, C.GTs[k]);
  }
}

std::string set_surface_vertices(Geometry *G, int ns, const double *V, int nV)
{
  if (ns < 0 || ns >= (int)G->Surfaces.size()) return "bad surface index";
  Surface &S = G->Surfaces[ns];
  if (nV != S.NumVertices()) return "vertex count differs from the mesh";
  for (long i = 0; i < 3L * nV; i++)
    if (!std::isfinite(V[i])) return "non-finite vertex coordinate";
  std::vector<double> old = S.V;
  S.V.assign(V, V + 3 * (size_t)nV);
  reinit_panels(&S);
  for (const Panel &P : S.Panels)
    if (!(P.Area > 0.0)) {
      S.V = old;
      reinit_panels(&S);
      return "a panel became degenerate";
    }
  S.MeshFile.clear();   // no longer the file's mesh: never a mate
  // edge data as build_surface derives it from the vertices
  for (Edge &E : S.Edges) {
    const double *VL = S.V.data() + 3 * E.iV1, *VG = S.V.data() + 3 * E.iV2;
    for (int i = 0; i < 3; i++) E.Centroid[i] = (VL[i] + VG[i]) / 2.0;
    E.Length = dist3(VL, VG);
    E.Radius = dist3(E.Centroid, S.V.data() + 3 * E.iQP);
    E.Radius = std::fmax(E.Radius, dist3(E.Centroid, VL));
    E.Radius = std::fmax(E.Radius, dist3(E.Centroid, VG));
    if (E.iQM >= 0) E.Radius = std::fmax(E.Radius, dist3(E.Centroid, S.V.data() + 3 * E.iQM));
  }
  if ((int)G->SurfaceMoved.size() != (int)G->Surfaces.size()) G->SurfaceMoved.assign(G->Surfaces.size(), 0);
  G->SurfaceMoved[ns] = 1;
  return "";
}

void untransform_geometry(Geometry *G)
{
  for (Surface &S : G->Surfaces) {
    S.GT.unapply(S.V.data(), S.V.size() / 3);
    for (Edge &E : S.Edges) S.GT.unapply(E.Centroid, 1);
    S.GT.reset();
    reinit_panels(&S);
  }
}

std::string read_scuffgeo(const std::string &fn, Geometry *G)
{
  std::ifstream in(fn);
  if (!in) return "could not open " + fn;
  std::string dir;
  size_t sl = fn.rfind('/');
  if (sl != std::string::npos) dir = fn.substr(0, sl + 1);
  std::string line, mesh, mat = "PEC", label;
  bool inobj = false, moved = false;
  GTransform T;
  int lineno = 0;
  // user-defined materials: MATERIAL name / constants / Eps(w) = ... /
  // Mu(w) = ... / ENDMATERIAL (UserDefinedMaterials.cc:145-320)
  std::map<std::string, Material> user;
  std::string inmat;
  std::map<std::string, double> consts;
  Material cur;
  while (std::getline(in, line)) {
    lineno++;
    std::string where = fn + ":" + std::to_string(lineno) + ": ";
    if (!inmat.empty()) {
      size_t hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
      size_t b = line.find_first_not_of(" \t\r\n");
      if (b == std::string::npos) continue;
      std::string body = line.substr(b);
      if (!strncasecmp(body.c_str(), "ENDMATERIAL", 11)) {
        user[inmat] = cur;
        inmat.clear();
        continue;
      }
      size_t eq = body.find('=');
      if (eq == std::string::npos) return where + "unrecognized input in MATERIAL " + inmat;
      std::string lhs = body.substr(0, eq), rhs = body.substr(eq + 1);
      while (!lhs.empty() && std::isspace((unsigned char)lhs.back())) lhs.pop_back();
      size_t semi = rhs.rfind(';');
      if (semi != std::string::npos) rhs.resize(semi);
      while (!rhs.empty() && std::isspace((unsigned char)rhs.back())) rhs.pop_back();
      if (!strncasecmp(lhs.c_str(), "Eps(w)", 6) || !strncasecmp(lhs.c_str(), "Mu(w)", 5)) {
        CExprPtr e;
        std::string err = parse_cexpr(rhs, consts, &e);
        if (!err.empty()) return where + "invalid expression: " + err;
        if (!strncasecmp(lhs.c_str(), "Eps(w)", 6)) cur.epsx = e; else cur.mux = e;
      } else {
        char *end = nullptr;
        const char *r0 = rhs.c_str();
        double v = std::strtod(r0, &end);
        if (end == r0) return where + "invalid constant value";
        while (*end && std::isspace((unsigned char)*end)) end++;
        if (*end) return where + "junk at end of line";
        consts[lhs] = v;
      }
      continue;
    }
    size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::stringstream ss(line);
    std::vector<std::string> tok;
    std::string w;
    while (ss >> w) tok.push_back(w);
    if (tok.empty()) continue;
    const std::string &kw = tok[0];
    if (!inobj && !strcasecmp(kw.c_str(), "MATERIAL")) {
      if (tok.size() != 2) return where + "no name given for MATERIAL";
      inmat = tok[1];
      consts.clear();
      cur = Material();
      cur.kind = MAT_EXPR;
      continue;
    }
    if (!strcasecmp(kw.c_str(), "OBJECT")) {
      if (inobj) return where + "OBJECT inside OBJECT";
      inobj = true; mesh.clear(); mat = "PEC"; moved = false; T = GTransform();
      label = tok.size() > 1 ? tok[1] : "";
    } else if (!inobj) {
      return where + "unsupported keyword " + kw + " (only OBJECT sections are handled)";
    } else if (!strcasecmp(kw.c_str(), "MESHFILE")) {
      if (tok.size() != 2) return where + "MESHFILE needs one file name";
      mesh = tok[1];
    } else if (!strcasecmp(kw.c_str(), "MATERIAL")) {
      if (tok.size() != 2) return where + "MATERIAL needs one name";
      mat = tok[1];
    } else if (!strcasecmp(kw.c_str(), "DISPLACED") || !strcasecmp(kw.c_str(), "DISP") ||
               !strcasecmp(kw.c_str(), "ROTATED") || !strcasecmp(kw.c_str(), "ROT")) {
      GTransform B;
      size_t used;
      std::string e = parse_transform(tok, 0, &B, &used);
      if (!e.empty()) return where + e;
      if (used != tok.size()) return where + "junk at end of line";
      if (moved) T.compose(B); else T = B;
      moved = true;
    } else if (!strcasecmp(kw.c_str(), "ENDOBJECT")) {
      if (mesh.empty()) return where + "OBJECT without MESHFILE";
      std::vector<double> V; std::vector<int> Tri;
      std::string err = read_gmsh(mesh[0] == '/' ? mesh : dir + mesh, &V, &Tri);
      if (!err.empty()) return err;
      if (moved) T.apply(V.data(), V.size() / 3);
      Material m; bool pec = false;
      auto um = user.find(mat);
      if (um != user.end()) m = um->second;
      else {
        err = parse_material(mat, &m, &pec);
        if (!err.empty()) return where + err;
      }
      Surface S;
      err = build_surface(V.data(), (int)V.size() / 3, Tri.data(), (int)Tri.size() / 3, pec, &S);
      if (!err.empty()) return where + err;
      S.MeshFile = mesh;
      S.MaterialName = pec ? std::string("PEC") : mat;
      S.Label = label.empty() ? mesh : label;
      add_surface(G, std::move(S), m);
      inobj = false;
    } else {
      return where + "unsupported keyword " + kw + " in OBJECT";
    }
  }
  if (!inmat.empty()) return fn + ": MATERIAL " + inmat + " without ENDMATERIAL";
  if (inobj) return fn + ": OBJECT without ENDOBJECT";
  if (G->Surfaces.empty()) return fn + ": no objects";
  return "";
}

// ---------------------------------------------------------------------
// icosphere: level 0 is the icosahedron; every level splits each
// triangle (a,b,c) into (a,ab,ca), (ab,b,bc), (ca,bc,c), (ab,bc,ca) with
// unit-sphere midpoints u=(ua+ub)/|ua+ub|.  Panel 4t+i is child i of
// panel t of the previous level, so neighbouring panels have nearby
// indices.
// ---------------------------------------------------------------------
void icosphere(int level, double radius, const double *center,
               std::vector<double> *Vout, std::vector<int> *Tout)
{
  static const int faces[20][3] = {
      {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
      {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
      {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
      {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {
      {-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0},
      {0, -1, phi}, {0, 1, phi}, {0, -1, -phi}, {0, 1, -phi},
      {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
  std::vector<double> U;
  std::vector<int> T;
  for (int i = 0; i < 12; i++) {
    double n = std::sqrt(base[i][0] * base[i][0] + base[i][1] * base[i][1] + base[i][2] * base[i][2]);
    for (int j = 0; j < 3; j++) U.push_back(base[i][j] / n);
  }
  for (int t = 0; t < 20; t++) for (int j = 0; j < 3; j++) T.push_back(faces[t][j]);
  for (int l = 0; l < level; l++) {
    std::unordered_map<long long, int> mid;
    mid.reserve(T.size());
    long long nVmax = 10LL * (1LL << (2 * level)) + 2;
    std::vector<int> T2(4 * T.size());
    size_t nt = T.size() / 3;
    for (size_t t = 0; t < nt; t++) {
      int v[3] = {T[3 * t], T[3 * t + 1], T[3 * t + 2]}, m[3];
      for (int e = 0; e < 3; e++) {
        int a = v[e], b = v[(e + 1) % 3];
        long long key = (long long)std::min(a, b) * nVmax + std::max(a, b);
        auto it = mid.find(key);
        if (it == mid.end()) {
          double s[3];
          for (int j = 0; j < 3; j++) s[j] = U[3 * a + j] + U[3 * b + j];
          double n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
          int idx = (int)(U.size() / 3);
          for (int j = 0; j < 3; j++) U.push_back(s[j] / n);
          mid.emplace(key, idx);
          m[e] = idx;
        } else m[e] = it->second;
      }
      int *c = T2.data() + 12 * t;
      c[0] = v[0]; c[1] = m[0]; c[2] = m[2];
      c[3] = m[0]; c[4] = v[1]; c[5] = m[1];
      c[6] = m[2]; c[7] = m[1]; c[8] = v[2];
      c[9] = m[0]; c[10] = m[1]; c[11] = m[2];
    }
    T.swap(T2);
  }
  Vout->resize(U.size());
  for (size_t i = 0; i < U.size() / 3; i++)
    for (int j = 0; j < 3; j++) (*Vout)[3 * i + j] = (center ? center[j] : 0.0) + radius * U[3 * i + j];
  *Tout = T;
}

std::string write_gmsh(const std::string &fn, const std::vector<double> &V, const std::vector<int> &T)
{
  FILE *f = std::fopen(fn.c_str(), "w");
  if (!f) return "could not write " + fn;
  size_t nV = V.size() / 3, nT = T.size() / 3;
  std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%zu\n", nV);
  for (size_t i = 0; i < nV; i++) std::fprintf(f, "%zu %.17g %.17g %.17g\n", i + 1, V[3 * i], V[3 * i + 1], V[3 * i + 2]);
  std::fprintf(f, "$EndNodes\n$Elements\n%zu\n", nT);
  for (size_t t = 0; t < nT; t++) std::fprintf(f, "%zu 2 2 0 0 %d %d %d\n", t + 1, T[3 * t] + 1, T[3 * t + 1] + 1, T[3 * t + 2] + 1);
  std::fprintf(f, "$EndElements\n");
  std::fclose(f);
  return "";
}

}  // namespace scuffx
