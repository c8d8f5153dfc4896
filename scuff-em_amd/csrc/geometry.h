// geometry.h -- host-side mesh topology for the BEM system path.
//
// RWG panels and edges with the reference's indexing
// (libs/libscuff/libscuff.h:100-146, RWGSurface.cc:55-97,
// InitEdgeList.cc:126-245), compact objects in vacuum with one interior
// region per non-PEC object (RWGGeometry.cc:494-651, :708-724), and the
// material models the workloads use (libs/libMatProp/libMatProp.cc:501).
#pragma once

#include <complex>
#include <string>
#include <vector>

namespace scuffx {

using cdouble = std::complex<double>;

struct Panel {
  int VI[3];
  int EI[3];
  double Centroid[3];
  double ZHat[3];
  double Radius;
  double Area;
};

struct Edge {
  int iV1, iV2, iQP, iQM;
  double Centroid[3];
  double Length, Radius;
  int iPPanel, iMPanel, PIndex, MIndex;
  int Index;
};

enum MaterialKind { MAT_VACUUM = 0, MAT_CONST = 1, MAT_DRUDE = 2, MAT_SIGMA = 3 };
struct Material {
  int kind = MAT_VACUUM;
  cdouble eps = 1.0, mu = 1.0;
  double wp = 0.0, gamma = 0.0;   // Drude: eps = 1 - wp^2/(w(w+i gamma))
  double sigma = 0.0;             // SIGMA model: eps = Re(eps) + i sigma/w
  void eps_mu(cdouble omega, cdouble *e, cdouble *m) const;
};

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
  int NumVertices() const { return (int)(V.size() / 3); }
  int NumPanels() const { return (int)Panels.size(); }
  int NumEdges() const { return (int)Edges.size(); }
};

struct Geometry {
  std::vector<Surface> Surfaces;
  std::vector<Material> Regions;     // Regions[0] is the exterior medium
  std::vector<int> BFIndexOffset;
  int TotalBFs = 0;
  Geometry() { Regions.push_back(Material()); }
};

// Builds a surface from vertex coordinates and triangle vertex indices.
// Returns an empty string on success, an error message otherwise.
std::string build_surface(const double *V, int nV, const int *T, int nT,
                          bool pec, Surface *S);
int add_surface(Geometry *G, Surface &&S, const Material &interior);
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
