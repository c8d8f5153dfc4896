// geometry.h -- host-side mesh topology for the BEM system path.
//
// RWG panels and edges with the reference's indexing
// (libs/libscuff/libscuff.h:100-146, RWGSurface.cc:55-97,
// InitEdgeList.cc:126-245), compact objects in vacuum with one interior
// region per non-PEC object (RWGGeometry.cc:494-651, :708-724), and the
// material models the workloads use (libs/libMatProp/libMatProp.cc:501).
#pragma once

#include <complex>
#include <map>
#include <string>
#include <vector>

#include "cexpr.h"

namespace scuffx {

using cdouble = std::complex<double>;

struct Panel {
  int VI[3];
  int EI[3];
  double Centroid[3];
  double ZHat[3];
  double Radius;
  double Area;
  bool ZHatFlipped = false;   // normal reversed to point into the exterior region
};

struct Edge {
  int iV1, iV2, iQP, iQM;
  double Centroid[3];
  double Length, Radius;
  int iPPanel, iMPanel, PIndex, MIndex;
  int Index;
};

enum MaterialKind { MAT_VACUUM = 0, MAT_CONST = 1, MAT_DRUDE = 2, MAT_SIGMA = 3, MAT_EXPR = 4 };
struct Material {
  int kind = MAT_VACUUM;
  cdouble eps = 1.0, mu = 1.0;
  double wp = 0.0, gamma = 0.0;   // Drude: eps = 1 - wp^2/(w(w+i gamma))
  double sigma = 0.0;             // SIGMA model: eps = Re(eps) + i sigma/w
  CExprPtr epsx, mux;             // user-defined Eps(w), Mu(w) of w in rad/s (null: 1)
  void eps_mu(cdouble omega, cdouble *e, cdouble *m) const;
};

// rigid transformation X' = M X + DX (GTransformation.cc:26-142,
// :234-315); rotations and compositions in the reference's operation order
struct GTransform {
  double M[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double DX[3] = {0, 0, 0};
  void reset() { *this = GTransform(); }
  void displace(const double d[3]);
  void rotate(const double axis[3], double degrees);
  void compose(const GTransform &B);          // this <- B o this (B after this)
  void apply(double *X, size_t npoints) const;   // in place
  void unapply(double *X, size_t npoints) const;
};
// "DISPLACED x y z" / "ROTATED t ABOUT x y z" (and DISP / ROT) at tok[0];
// applies it to *out; *consumed = tokens used (GTransformation.cc:159-245)
std::string parse_transform(const std::vector<std::string> &tok, size_t first, GTransform *out, size_t *consumed);

// one named set of surface transformations (GTComplex, GTransformation.h)
struct GTComplex {
  std::string Tag;
  std::vector<std::string> Labels;   // empty label: none was given
  std::vector<GTransform> GTs;
};
// .trans file: TRANSFORMATION tag ... ENDTRANSFORMATION sections and
// one-line TRANS tag OBJECT a DISP ... entries (GTransformation.cc:373-580);
// an empty file name gives one identity complex tagged DEFAULT
std::string read_trans_file(const std::string &fn, std::vector<GTComplex> *out);

struct Surface {
  std::vector<double> V;        // 3*NumVertices
  std::vector<Panel> Panels;
  std::vector<Edge> Edges;      // interior edges, Edges[i].Index == i
  int NumExteriorEdges = 0;
  bool IsPEC = true;
  int RegionIndices[2] = {0, -1};
  int NumBFs = 0;
  std::string MeshFile;         // as named by the caller; empty for raw arrays
  std::string MaterialName;     // interior material as given ("PEC" for PEC)
  std::string Label;            // OBJECT label, mesh file name, or ByHand
  GTransform GT;                // transformations since the last UnTransform
  int NumVertices() const { return (int)(V.size() / 3); }
  int NumPanels() const { return (int)Panels.size(); }
  int NumEdges() const { return (int)Edges.size(); }
};

struct Geometry {
  std::vector<Surface> Surfaces;
  std::vector<Material> Regions;     // Regions[0] is the exterior medium
  std::vector<int> BFIndexOffset;
  std::vector<int> SurfaceMoved;     // set by the last transform_geometry
  int TotalBFs = 0;
  Geometry() { Regions.push_back(Material()); }
};

// index of the surface with this label (case-insensitive), or -1
int surface_by_label(const Geometry &G, const std::string &label);
// empty if every transformation of C names an existing surface
// (RWGGeometry::CheckGTCList, RWGGeometry.cc:931-943)
std::string check_gtc(const Geometry &G, const GTComplex &C);
// RWGGeometry::Transform / UnTransform (RWGGeometry.cc:888-921,
// RWGSurface.cc:588-667): vertices and edge centroids are transformed,
// panels re-initialized from the vertices; UnTransform applies the
// inverse of the accumulated transformation to the current coordinates
void transform_geometry(Geometry *G, const GTComplex &C);
// one surface, by index (the Transform of a single labelled entry)
void transform_surface(Geometry *G, int ns, const GTransform &D);
void untransform_geometry(Geometry *G);
// new coordinates for the vertices of surface ns (same mesh topology):
// panels and edges are re-derived from them, the normals keep their
// orientation; an empty string on success
std::string set_surface_vertices(Geometry *G, int ns, const double *V, int nV);

// Builds a surface from vertex coordinates and triangle vertex indices.
// Returns an empty string on success, an error message otherwise.
std::string build_surface(const double *V, int nV, const int *T, int nT,
                          bool pec, Surface *S);
int add_surface(Geometry *G, Surface &&S, const Material &interior);
// panel normals made to point into each surface's exterior region
// (RWGGeometry.cc:653-690); called by add_surface for every surface
void orient_normals(Geometry *G);
// Mate[ns] = the first earlier surface identical to ns, or -1: same mesh
// file, same exterior medium, both PEC or the same interior material
// (RWGGeometry.cc:733-770; MESHTAG is not supported, so tags always agree)
std::vector<int> find_mates(const Geometry &G);

// GMSH 2.x / legacy reader (ParseMeshFiles.cc:69-259).
std::string read_gmsh(const std::string &fn, std::vector<double> *V,
                      std::vector<int> *T);
// minimal .scuffgeo reader: OBJECT / MESHFILE / MATERIAL / ENDOBJECT
// (RWGGeometry.cc:433-571, RWGSurface.cc:119-217)
std::string read_scuffgeo(const std::string &fn, Geometry *G);
// recursive-subdivision icosphere used to generate the benchmark meshes
void icosphere(int level, double radius, const double *center,
               std::vector<double> *V, std::vector<int> *T);
std::string write_gmsh(const std::string &fn, const std::vector<double> &V,
                       const std::vector<int> &T);
// parses "PEC", "CONST_EPS_<re>[+/-<im>i]", "DRUDE_<wp>_<gamma>"
std::string parse_material(const std::string &s, Material *m, bool *pec);

}  // namespace scuffx
