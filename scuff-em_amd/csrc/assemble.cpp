// assemble.cpp -- host orchestration of the BEM matrix assembly.
//
// Call structure restated from the reference:
//   AssembleBEMMatrix (AssembleBEMMatrix.cc:684-773): surface pairs
//   ns <= nsp, each block filled by GetSurfaceSurfaceInteractions
//   (SurfaceSurfaceInteractions.cc:426-585), then M(r,c) = M(c,r) below
//   the diagonal.  Inside a block every RWG pair sums four panel-pair
//   integrals (EdgeEdgeInteractions.cc:154-188) whose algorithm is
//   chosen per panel pair (PanelPanelInteractions.cc:357-561).
// Here each class of panel pair is a separate kernel: far pairs (fixed
// rule), near pairs (desingularized rule + cached singular terms), self
// pairs (Taylor-Duffy); a final kernel combines them per edge pair.
#include "assemble.h"
#include "fippi_cache.h"
#include "prof.h"
#include "ws.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace scuffx {

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { err_ = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x; return -3; } } while (0)

static double dist3h(const double *a, const double *b)
{
  double d = 0;
  for (int i = 0; i < 3; i++) d += (a[i] - b[i]) * (a[i] - b[i]);
  return std::sqrt(d);
}

// relative centroid distance and common-vertex count with the common
// vertices moved to the front (AssessPanelPair.cc:52-172).  ia/ib
// receive the mesh-local vertex slots in call order.
int assess_pair(const Surface &Sa, int pa, const Surface &Sb, int pb, double *rRel, int ia[3], int ib[3])
{
  const Panel &Pa = Sa.Panels[pa], &Pb = Sb.Panels[pb];
  for (int i = 0; i < 3; i++) { ia[i] = i; ib[i] = i; }
  double DC, r2, rMax = std::fmax(Pa.Radius, Pb.Radius);
  DC = Pa.Centroid[0] - Pb.Centroid[0]; r2 = DC * DC;
  DC = Pa.Centroid[1] - Pb.Centroid[1]; r2 += DC * DC;
  DC = Pa.Centroid[2] - Pb.Centroid[2]; r2 += DC * DC;
  *rRel = std::sqrt(r2) / rMax;
  if (*rRel > 2.0) return 0;
  double thr = std::fabs(1.0e-6 * rMax);
  int ncv = 0, CVa[3], CVb[3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double *x = Sa.V.data() + 3 * Pa.VI[a], *y = Sb.V.data() + 3 * Pb.VI[b];
      if (std::fabs(x[0] - y[0]) + std::fabs(x[1] - y[1]) + std::fabs(x[2] - y[2]) <= 3 * thr) {
        CVa[ncv] = a; CVb[ncv] = b; ncv++;
      }
    }
  if (ncv == 1) {
    for (int i = 0; i < 3; i++) { ia[i] = (CVa[0] + i) % 3; ib[i] = (CVb[0] + i) % 3; }
  } else if (ncv == 2) {
    ia[0] = CVa[0]; ia[1] = CVa[1]; ia[2] = 3 - CVa[0] - CVa[1];
    ib[0] = CVb[0]; ib[1] = CVb[1]; ib[2] = 3 - CVb[0] - CVb[1];
  } else if (ncv == 3) {
    for (int i = 0; i < 3; i++) { ia[i] = CVa[i]; ib[i] = CVb[i]; }
  }
  return ncv;
}

// vertex order with 1e-6 relative coordinate tolerance (AssessPanelPair.cc:250-267)
static int vlt(const double *V1, const double *V2)
{
  double DV;
  DV = V1[0] - V2[0]; if (std::fabs(DV) > 1.0e-6 * std::fabs(V1[0])) return DV < 0.0;
  DV = V1[1] - V2[1]; if (std::fabs(DV) > 1.0e-6 * std::fabs(V1[1])) return DV < 0.0;
  DV = V1[2] - V2[2]; if (std::fabs(DV) > 1.0e-6 * std::fabs(V1[2])) return DV < 0.0;
  return 1;
}

// canonical ordering for the FIPPI key (AssessPanelPair.cc:342-455).
// For ncv == 0 the reference compares Va[iMina] with Va[iMinb] and, in
// the swapped case, takes OVa from Vb with panel a's sort indices; that
// behaviour is kept because it decides which cached record a pair uses.
static int canonical(const double *Va[3], const double *Vb[3], int ncv, const double *OVa[3], const double *OVb[3])
{
  int iMina, iMeda = 0, iMaxa, iMinb, iMaxb;
  if (ncv == 3 || ncv == 0) {
    if (vlt(Va[0], Va[1])) { iMina = 0; iMaxa = 1; } else { iMina = 1; iMaxa = 0; }
    if (vlt(Va[2], Va[iMina])) { iMeda = iMina; iMina = 2; }
    else if (vlt(Va[2], Va[iMaxa])) iMeda = 2;
    else { iMeda = iMaxa; iMaxa = 2; }
    if (ncv == 3) {
      OVa[0] = Va[iMina]; OVa[1] = Va[iMeda]; OVa[2] = Va[iMaxa];
      OVb[0] = Vb[iMina]; OVb[1] = Vb[iMeda]; OVb[2] = Vb[iMaxa];
      return 0;
    }
    if (vlt(Vb[0], Vb[1])) { iMinb = 0; iMaxb = 1; } else { iMinb = 1; iMaxb = 0; }
    if (vlt(Vb[2], Vb[iMinb])) iMinb = 2;
    else if (vlt(Vb[2], Vb[iMaxb])) { }
    else iMaxb = 2;
    if (vlt(Va[iMina], Va[iMinb])) {
      OVa[0] = Va[iMina]; OVa[1] = Va[iMeda]; OVa[2] = Va[iMaxa];
      OVb[0] = Vb[iMina]; OVb[1] = Vb[iMeda]; OVb[2] = Vb[iMaxa];
      return 0;
    }
    OVb[0] = Va[iMina]; OVb[1] = Va[iMeda]; OVb[2] = Va[iMaxa];
    OVa[0] = Vb[iMina]; OVa[1] = Vb[iMeda]; OVa[2] = Vb[iMaxa];
    return 1;
  }
  if (ncv == 2) {
    int a01 = vlt(Va[0], Va[1]) ? 0 : 1;
    if (vlt(Va[2], Vb[2])) {
      OVa[0] = Va[a01]; OVa[1] = Va[1 - a01]; OVa[2] = Va[2];
      OVb[0] = Vb[a01]; OVb[1] = Vb[1 - a01]; OVb[2] = Vb[2];
      return 0;
    }
    OVa[0] = Vb[a01]; OVa[1] = Vb[1 - a01]; OVa[2] = Vb[2];
    OVb[0] = Va[a01]; OVb[1] = Va[1 - a01]; OVb[2] = Va[2];
    return 1;
  }
  int a12 = vlt(Va[1], Va[2]) ? 1 : 2, b12 = vlt(Vb[1], Vb[2]) ? 1 : 2;
  if (vlt(Va[a12], Vb[b12])) {
    OVa[0] = Va[0]; OVa[1] = Va[a12]; OVa[2] = Va[3 - a12];
    OVb[0] = Vb[0]; OVb[1] = Vb[b12]; OVb[2] = Vb[3 - b12];
    return 0;
  }
  OVa[0] = Vb[0]; OVa[1] = Vb[b12]; OVa[2] = Vb[3 - b12];
  OVb[0] = Va[0]; OVb[1] = Va[a12]; OVb[2] = Va[3 - a12];
  return 1;
}


Assembler::Assembler(const Geometry &G) : G_(G)
{
  static std::atomic<unsigned long long> counter{0};
  serial_ = ++counter;
  ds_.resize(G.Surfaces.size());
  for (size_t s = 0; s < G.Surfaces.size(); s++) {
    const Surface &S = G.Surfaces[s];
    DevSurface &D = ds_[s];
    D.nV = S.NumVertices(); D.nP = S.NumPanels(); D.nE = S.NumEdges();
    std::vector<int> pvi(3 * D.nP), ei(4 * D.nE);
    std::vector<double> pc(4 * D.nP), el(D.nE);
    for (int p = 0; p < D.nP; p++) {
      for (int j = 0; j < 3; j++) { pvi[3 * p + j] = S.Panels[p].VI[j]; pc[4 * p + j] = S.Panels[p].Centroid[j]; }
      pc[4 * p + 3] = S.Panels[p].Radius;
    }
    for (int e = 0; e < D.nE; e++) {
      const Edge &E = S.Edges[e];
      ei[4 * e] = E.iPPanel; ei[4 * e + 1] = E.iMPanel; ei[4 * e + 2] = E.PIndex; ei[4 * e + 3] = E.MIndex;
      el[e] = E.Length;
    }
    if (hipMalloc(&D.V, sizeof(double) * 3 * D.nV) != hipSuccess ||
        hipMalloc(&D.PVI, sizeof(int) * 3 * D.nP) != hipSuccess ||
        hipMalloc(&D.PC, sizeof(double) * 4 * D.nP) != hipSuccess ||
        hipMalloc(&D.EI, sizeof(int) * 4 * std::max(D.nE, 1)) != hipSuccess ||
        hipMalloc(&D.EL, sizeof(double) * std::max(D.nE, 1)) != hipSuccess) {
      err_ = "device allocation failed for the geometry";
      return;
    }
    hipMemcpy(D.V, S.V.data(), sizeof(double) * 3 * D.nV, hipMemcpyHostToDevice);
    hipMemcpy(D.PVI, pvi.data(), sizeof(int) * 3 * D.nP, hipMemcpyHostToDevice);
    hipMemcpy(D.PC, pc.data(), sizeof(double) * 4 * D.nP, hipMemcpyHostToDevice);
    if (D.nE) {
      hipMemcpy(D.EI, ei.data(), sizeof(int) * 4 * D.nE, hipMemcpyHostToDevice);
      hipMemcpy(D.EL, el.data(), sizeof(double) * D.nE, hipMemcpyHostToDevice);
    }
  }
  hipMalloc(&d_regs_, sizeof(RegionK) * 4);
  mates_ = find_mates(G);
  build_near();
}

Assembler::~Assembler()
{
  if (hip_down()) return;   // after scuffx_shutdown: left to the process exit
  for (auto &D : ds_) { hipFree(D.V); hipFree(D.PVI); hipFree(D.PC); hipFree(D.EI); hipFree(D.EL); }
  for (auto &B : blocks_) {
    hipFree(B.d_rowPtr); hipFree(B.d_col); hipFree(B.d_idx); hipFree(B.d_pairs);
    hipFree(B.bands.d_cols); hipFree(B.bands.d_colSlots); hipFree(B.bands.d_slots);
  }
  for (auto &e : evpool_) hipEventDestroy(e);
  for (void *p : vg_alloc_) hipFree(p);
  hipFree(d_jobs_); hipFree(d_records_);
  hipFree(d_HP_); hipFree(d_HN_); hipFree(d_HS_); hipFree(d_scratch_); hipFree(d_stage_);
  hipFree(d_regs_);
  hipFree(d_swjobs_);
  hipFree(d_regtab_);
}

int Assembler::ensure(double **p, size_t *cap, size_t need)
{
  if (need <= *cap && *p) return 0;
  alloc_gen_++;
  if (*p) hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (need == 0) need = 1;
  if (hipMalloc(p, need * sizeof(double)) != hipSuccess) { *p = nullptr; err_ = "device allocation failed"; return -2; }
  *cap = need;
  return 0;
}

// Near pairs of every surface-pair block, the order in which the
// reference's serial edge-pair traversal first touches each ordered
// pair, and the deduplicated FIPPI job list (first touch wins, as in
// FIPPICache::GetQIFIPPIData, FIPPICache.cc:135-187).
void Assembler::build_near()
{
  const int NS = (int)G_.Surfaces.size();
  struct Occ { uint64_t pos; int blk, pa, pb; };
  std::vector<Occ> occ;
  int bi = 0;
  for (int sa = 0; sa < NS; sa++)
    for (int sb = sa; sb < NS; sb++, bi++) {
      const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
      NearBlock B;
      B.sa = sa; B.sb = sb;
      const int nPa = Sa.NumPanels(), nPb = Sb.NumPanels();
      // uniform grid on panel centroids, cell = 4 x largest radius
      double Rmax = 0;
      for (auto &P : Sa.Panels) Rmax = std::max(Rmax, P.Radius);
      for (auto &P : Sb.Panels) Rmax = std::max(Rmax, P.Radius);
      double cell = 4.0 * Rmax * 1.000001 + 1e-300;
      double lo[3] = {1e300, 1e300, 1e300};
      for (auto &P : Sb.Panels) for (int j = 0; j < 3; j++) lo[j] = std::min(lo[j], P.Centroid[j]);
      auto cidx = [&](const double *c, int j) { return (long)std::floor((c[j] - lo[j]) / cell); };
      std::unordered_map<long long, std::vector<int>> grid;
      auto gkey = [](long x, long y, long z) { return ((long long)(x + 1048576) << 42) ^ ((long long)(y + 1048576) << 21) ^ (long long)(z + 1048576); };
      for (int p = 0; p < nPb; p++) {
        const double *c = Sb.Panels[p].Centroid;
        grid[gkey(cidx(c, 0), cidx(c, 1), cidx(c, 2))].push_back(p);
      }
      B.rowPtr.assign(nPa + 1, 0);
      std::vector<std::vector<int>> rows(nPa);
      for (int pa = 0; pa < nPa; pa++) {
        const double *c = Sa.Panels[pa].Centroid;
        long cx = cidx(c, 0), cy = cidx(c, 1), cz = cidx(c, 2);
        for (long dx = -1; dx <= 1; dx++) for (long dy = -1; dy <= 1; dy++) for (long dz = -1; dz <= 1; dz++) {
          auto it = grid.find(gkey(cx + dx, cy + dy, cz + dz));
          if (it == grid.end()) continue;
          for (int pb : it->second) {
            if (sa == sb && pb == pa) continue;
            int ia[3], ib[3];
            double rRel;
            int ncv = assess_pair(Sa, pa, Sb, pb, &rRel, ia, ib);
            if (rRel > 4.0) continue;
            if (ncv == 3) continue;   // coincident distinct panels: not supported
            rows[pa].push_back(pb);
          }
        }
        std::sort(rows[pa].begin(), rows[pa].end());
      }
      for (int pa = 0; pa < nPa; pa++) B.rowPtr[pa + 1] = B.rowPtr[pa] + (int)rows[pa].size();
      B.col.resize(B.rowPtr[nPa]);
      B.idx.assign(B.rowPtr[nPa], -1);
      for (int pa = 0; pa < nPa; pa++) std::copy(rows[pa].begin(), rows[pa].end(), B.col.begin() + B.rowPtr[pa]);
      blocks_.push_back(B);
      // first touch of each ordered pair in the (ea, eb>=ea|0, PP/PM/MP/MM) walk
      NearBlock &BB = blocks_.back();
      const int nEa = Sa.NumEdges(), nEb = Sb.NumEdges();
      std::vector<uint64_t> first(BB.col.size(), ~0ULL);
      for (int ea = 0; ea < nEa; ea++) {
        const Edge &Ea = Sa.Edges[ea];
        int pas[2] = {Ea.iPPanel, Ea.iMPanel};
        for (int s = 0; s < 2; s++) {
          int pa = pas[s];
          if (pa < 0) continue;
          for (int q = BB.rowPtr[pa]; q < BB.rowPtr[pa + 1]; q++) {
            int pb = BB.col[q];
            for (int le = 0; le < 3; le++) {
              int eb = Sb.Panels[pb].EI[le];
              if (eb < 0 || eb >= nEb) continue;
              if (Sb.Edges[eb].iPPanel != pb && Sb.Edges[eb].iMPanel != pb) continue;
              if (sa == sb && eb < ea) continue;
              int t = (Sb.Edges[eb].iPPanel == pb) ? 0 : 1;
              uint64_t pos = ((uint64_t)ea * (uint64_t)nEb + (uint64_t)eb) * 4 + (uint64_t)(2 * s + t);
              if (pos < first[q]) first[q] = pos;
            }
          }
        }
      }
      // ordered near pairs that no edge pair of the block reaches (for
      // example (pb, pa) when every edge of pb precedes every edge of pa
      // on one surface) get no record: idx stays -1
      for (size_t q = 0; q < BB.col.size(); q++) {
        if (first[q] == ~0ULL) continue;
        int pa = (int)(std::upper_bound(BB.rowPtr.begin(), BB.rowPtr.end(), (int)q) - BB.rowPtr.begin()) - 1;
        occ.push_back({first[q], bi, pa, (int)q});
      }
    }
  // global first-use order: blocks in order, then traversal position
  std::stable_sort(occ.begin(), occ.end(), [](const Occ &x, const Occ &y) {
    if (x.blk != y.blk) return x.blk < y.blk;
    return x.pos < y.pos;
  });
  std::unordered_map<Key15, int, Key15Hash> keymap;
  for (auto &B : blocks_) B.pairs.clear();
  for (const Occ &o : occ) {
    NearBlock &B = blocks_[o.blk];
    const Surface &Sa = G_.Surfaces[B.sa], &Sb = G_.Surfaces[B.sb];
    int q = o.pb;   // CSR slot
    int pa = o.pa, pb = B.col[q];
    int ia[3], ib[3];
    double rRel;
    int ncv = assess_pair(Sa, pa, Sb, pb, &rRel, ia, ib);
    NearPair P;
    std::memset(&P, 0, sizeof(P));
    const double *Va[3], *Vb[3];
    for (int i = 0; i < 3; i++) {
      Va[i] = Sa.V.data() + 3 * Sa.Panels[pa].VI[ia[i]];
      Vb[i] = Sb.V.data() + 3 * Sb.Panels[pb].VI[ib[i]];
      for (int c = 0; c < 3; c++) {
        P.Va[3 * i + c] = Va[i][c];
        P.Vb[3 * i + c] = Vb[i][c];
        P.Qa[3 * i + c] = Sa.V[3 * Sa.Panels[pa].VI[i] + c];
        P.Qb[3 * i + c] = Sb.V[3 * Sb.Panels[pb].VI[i] + c];
      }
    }
    const double *OVa[3], *OVb[3];
    int fl = canonical(Va, Vb, ncv, OVa, OVb);
    for (int c = 0; c < 3; c++) {
      P.OVa0[c] = OVa[0][c]; P.OVa1[c] = OVa[1][c]; P.OVa2[c] = OVa[2][c];
      P.OVb0[c] = OVb[0][c]; P.OVb1[c] = OVb[1][c]; P.OVb2[c] = OVb[2][c];
    }
    P.flipped = fl;
    P.ncv = ncv;
    const Key15 K = fippi_key(OVa, OVb);
    auto it = keymap.find(K);
    if (it == keymap.end()) {
      FippiJob J;
      std::memset(&J, 0, sizeof(J));
      for (int i = 0; i < 3; i++) for (int c = 0; c < 3; c++) { J.OVa[3 * i + c] = OVa[i][c]; J.OVb[3 * i + c] = OVb[i][c]; }
      J.ncv = ncv;
      int id = (int)jobs_.size();
      jobs_.push_back(J);
      job_keys_.push_back(K);
      keymap.emplace(K, id);
      P.rec = id;
    } else P.rec = it->second;
    B.idx[q] = (int)B.pairs.size();
    B.pairs.push_back(P);
  }
  for (auto &B : blocks_) {
    hipMalloc(&B.d_rowPtr, sizeof(int) * B.rowPtr.size());
    hipMalloc(&B.d_col, sizeof(int) * std::max<size_t>(1, B.col.size()));
    hipMalloc(&B.d_idx, sizeof(int) * std::max<size_t>(1, B.idx.size()));
    hipMalloc(&B.d_pairs, sizeof(NearPair) * std::max<size_t>(1, B.pairs.size()));
    hipMemcpy(B.d_rowPtr, B.rowPtr.data(), sizeof(int) * B.rowPtr.size(), hipMemcpyHostToDevice);
    if (!B.col.empty()) {
      hipMemcpy(B.d_col, B.col.data(), sizeof(int) * B.col.size(), hipMemcpyHostToDevice);
      hipMemcpy(B.d_idx, B.idx.data(), sizeof(int) * B.idx.size(), hipMemcpyHostToDevice);
      hipMemcpy(B.d_pairs, B.pairs.data(), sizeof(NearPair) * B.pairs.size(), hipMemcpyHostToDevice);
    }
    build_bands(B);
  }
  stats_.nfippi = (long)jobs_.size();
}

// Column bands of a block.  The key of a column edge is its higher panel
// index; edges are numbered at the second visit of a panel scan
// (InitEdgeList.cc:163-167), so on one surface the key grows with the edge
// index and a band is a run of consecutive columns.  The band width `ch`
// keeps one band's far records near SCUFFX_FAR_BAND_MB (default 1024 MB):
// records of 2 regions x 9 combinations x {H0, H1}, and about 1.3 slots per
// panel of the key range (the lower panels of the band's edges).
void Assembler::build_bands(NearBlock &B)
{
  FarBands &F = B.bands;
  const Surface &Sa = G_.Surfaces[B.sa], &Sb = G_.Surfaces[B.sb];
  const int nPa = Sa.NumPanels(), nPb = Sb.NumPanels(), nEa = Sa.NumEdges(), nEb = Sb.NumEdges();
  const bool same = (B.sa == B.sb);
  double budget = 1024.0 * 1048576.0;
  if (const char *e = getenv("SCUFFX_FAR_BAND_MB")) budget = std::max(1.0, atof(e)) * 1048576.0;
  const double rec_bytes = 2 * 9 * 4 * sizeof(double);
  long ch = (long)(budget / (1.3 * std::max(nPa, 1) * rec_bytes));
  ch = std::max(32L, std::min(ch, (long)std::max(nPb, 1)));
  F.ch = (int)ch;
  F.rowsCap = std::max(nPa, 1);
  auto key = [&](int e) { return std::max(Sb.Edges[e].iPPanel, Sb.Edges[e].iMPanel); };
  const int nbands = (int)((nPb + ch - 1) / ch);
  std::vector<std::vector<int>> bcols(std::max(nbands, 1));
  for (int e = 0; e < nEb; e++) {
    int k = key(e);
    if (k < 0) continue;
    bcols[k / ch].push_back(e);
  }
  // row panels of each row-edge prefix (diagonal blocks): 1 + the highest
  // panel index among edges 0..ea
  std::vector<int> prefRow(nEa + 1, 0);
  for (int e = 0; e < nEa; e++)
    prefRow[e + 1] = std::max(prefRow[e], 1 + std::max(Sa.Edges[e].iPPanel, Sa.Edges[e].iMPanel));
  // near pairs by column panel (pa values of the CSR entries (pa, pb))
  std::vector<std::vector<int>> nearByCol(nPb);
  for (int pa = 0; pa < nPa; pa++)
    for (int q = B.rowPtr[pa]; q < B.rowPtr[pa + 1]; q++) nearByCol[B.col[q]].push_back(pa);
  F.uniqueFar = same ? (long)nPa * (nPa - 1) / 2 - (long)B.col.size() / 2 : (long)nPa * nPb - (long)B.col.size();
  F.colStart.assign(1, 0);
  F.slotStart.assign(1, 0);
  for (int k = 0; k < nbands; k++) {
    const std::vector<int> &cl = bcols[k];
    if (cl.empty()) continue;
    std::vector<int> sl;
    for (int e : cl) {
      if (Sb.Edges[e].iPPanel >= 0) sl.push_back(Sb.Edges[e].iPPanel);
      if (Sb.Edges[e].iMPanel >= 0) sl.push_back(Sb.Edges[e].iMPanel);
    }
    std::sort(sl.begin(), sl.end());
    sl.erase(std::unique(sl.begin(), sl.end()), sl.end());
    int maxe = 0;
    for (int e : cl) {
      F.cols.push_back(e);
      const int pb2[2] = {Sb.Edges[e].iPPanel, Sb.Edges[e].iMPanel};
      for (int t = 0; t < 2; t++)
        F.colSlots.push_back(pb2[t] < 0 ? -1 : (int)(std::lower_bound(sl.begin(), sl.end(), pb2[t]) - sl.begin()));
      maxe = std::max(maxe, e);
    }
    const int rowEdges = same ? maxe + 1 : nEa;
    const int rowLimit = same ? prefRow[std::min(rowEdges, nEa)] : nPa;
    long fp = 0;
    for (int pb : sl) {
      long n = rowLimit;
      if (same && pb < rowLimit) n--;
      for (int pa : nearByCol[pb]) if (pa < rowLimit) n--;
      fp += n;
    }
    F.slots.insert(F.slots.end(), sl.begin(), sl.end());
    F.colStart.push_back((int)F.cols.size());
    F.slotStart.push_back((int)F.slots.size());
    F.rowEdges.push_back(rowEdges);
    F.rowLimit.push_back(rowLimit);
    F.farPairs.push_back(fp);
    F.maxSlots = std::max(F.maxSlots, (int)sl.size());
  }
  auto up = [](int **d, const std::vector<int> &h) {
    hipMalloc(d, sizeof(int) * std::max<size_t>(1, h.size()));
    if (!h.empty()) hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice);
  };
  up(&F.d_cols, F.cols);
  up(&F.d_colSlots, F.colSlots);
  up(&F.d_slots, F.slots);
}

int Assembler::far_and_stamp(NearBlock &B, const DevSurface &Da, const DevSurface &Db, const Surface &Sa,
                             const Surface &Sb, int nreg, const RegionK *dreg, int needH1, MatOut M, long rowOff,
                             long colOff, hipStream_t stream, double *t_far, double *t_stamp)
{
  const FarBands &F = B.bands;
  const int same = (B.sa == B.sb);
  const int stride = needH1 ? 4 : 2;
  const long recd = (long)std::max(nreg, 1) * 9 * stride;
  if (ensure(&d_HP_, &cap_HP_, (size_t)std::max(F.maxSlots, 1) * F.rowsCap * recd)) return -2;
  const bool timed = t_far || t_stamp;
  const int nb = F.count();
  if (timed) {
    while ((int)evpool_.size() < 2 * nb + 1) {
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return -3;
      evpool_.push_back(e);
    }
    hipEventRecord(evpool_[0], stream);
  }
  // algorithmic flops per (panel pair, region): 36 point pairs of 98
  // (PMCHWT) or 56 (EFIE) flops plus the 9-combination epilogue; the
  // pairs a band recomputes (its lower slot panels) are not counted
  const double wpair = needH1 ? 3780.0 : 2160.0;
  long computed = 0;
  for (long v : F.farPairs) computed += v;
  const double useful = computed > 0 ? (double)F.uniqueFar / (double)computed : 0.0;
  const double nbf_col = Sb.IsPEC ? 1.0 : 2.0, nbf_row = Sa.IsPEC ? 1.0 : 2.0;
  for (int k = 0; k < nb; k++) {
    const int c0 = F.colStart[k], nc = F.colStart[k + 1] - c0;
    const int s0 = F.slotStart[k], ns = F.slotStart[k + 1] - s0;
    if (nreg > 0) {
      int tok = prof_begin(stream, PK_FAR);
      HIPCHK(launch_far(stream, Da.V, Da.PVI, Da.PC, F.rowLimit[k], Db.V, Db.PVI, Db.PC, F.d_slots + s0, ns, same,
                        B.d_rowPtr, B.d_col, nreg, dreg, needH1, d_HP_, F.rowsCap, far_bits_ == 32));
      prof_end(stream, tok, wpair * (double)F.farPairs[k] * useful * nreg);
    }
    if (timed) hipEventRecord(evpool_[2 * k + 1], stream);
    {
      // matrix bytes the band writes (upper triangle of a diagonal block)
      double entries = 0;
      for (int j = 0; j < nc; j++) {
        const int eb = F.cols[c0 + j];
        entries += same ? (double)(eb + 1) : (double)Sa.NumEdges();
      }
      entries *= nbf_row * nbf_col;
      int tok = prof_begin(stream, PK_STAMP);
      HIPCHK(launch_stamp(stream, Da.nE, Da.EI, Da.EL, F.d_cols + c0, F.d_colSlots + 2 * c0, nc, F.rowEdges[k],
                          Db.EI, Db.EL, same, Sa.IsPEC, Sb.IsPEC, nreg, dreg, needH1, d_HP_, F.rowsCap, d_HN_, d_HS_,
                          B.d_rowPtr, B.d_col, B.d_idx, M, (int)rowOff, (int)colOff));
      prof_end(stream, tok, 16.0 * entries);
    }
    if (timed) hipEventRecord(evpool_[2 * k + 2], stream);
  }
  if (timed && nb > 0) {
    HIPCHK(hipEventSynchronize(evpool_[2 * nb]));
    for (int k = 0; k < nb; k++) {
      float ms = 0;
      if (t_far && hipEventElapsedTime(&ms, evpool_[2 * k], evpool_[2 * k + 1]) == hipSuccess) *t_far += ms;
      if (t_stamp && hipEventElapsedTime(&ms, evpool_[2 * k + 1], evpool_[2 * k + 2]) == hipSuccess) *t_stamp += ms;
    }
  }
  return 0;
}

int Assembler::build_fippi(hipStream_t stream)
{
  if (fippi_ready_) return 0;
  if (!err_.empty()) return -3;
  const int nj = (int)jobs_.size();
  auto t0 = std::chrono::steady_clock::now();
  // records already in the process-wide cache (preloaded from a file or
  // computed for an earlier geometry) are copied, the others computed
  std::vector<double> rec(std::max(nj, 1) * (size_t)SCUFFX_QIF_LEN, 0.0);
  std::vector<int> todo;
  for (int j = 0; j < nj; j++)
    if (!fippi_cache_find(job_keys_[j], rec.data() + (size_t)j * SCUFFX_QIF_LEN)) todo.push_back(j);
  stats_.nfippi_cached = nj - (long)todo.size();
  if (nj > 0) {
    HIPCHK(hipMalloc(&d_jobs_, sizeof(FippiJob) * nj));
    HIPCHK(hipMalloc(&d_records_, sizeof(double) * SCUFFX_QIF_LEN * nj));
    HIPCHK(hipMemcpyAsync(d_jobs_, jobs_.data(), sizeof(FippiJob) * nj, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpy(d_records_, rec.data(), sizeof(double) * SCUFFX_QIF_LEN * nj, hipMemcpyHostToDevice));
  }
  const int nt = (int)todo.size();
  if (nt > 0) {
    int *d_counter = nullptr, *d_ovf = nullptr;
    HIPCHK(hipMalloc(&d_counter, sizeof(int)));
    HIPCHK(hipMalloc(&d_ovf, sizeof(int) * nj));
    HIPCHK(hipMemsetAsync(d_ovf, 0, sizeof(int) * nj, stream));
    HIPCHK(hipMemsetAsync(d_counter, 0, sizeof(int), stream));
    // pass 1: small heaps (4096 regions) for every job to compute, 2048 waves
    int cap1 = 4096, slots1 = std::min(2048, std::max(1, nt));
    double *hd = nullptr, *hk = nullptr;
    int *hs = nullptr, *d_todo = nullptr;
    if (nt < nj) {
      HIPCHK(hipMalloc(&d_todo, sizeof(int) * nt));
      HIPCHK(hipMemcpy(d_todo, todo.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&hd, sizeof(double) * (size_t)cap1 * 76 * slots1));
    HIPCHK(hipMalloc(&hk, sizeof(double) * (size_t)cap1 * slots1));
    HIPCHK(hipMalloc(&hs, sizeof(int) * (size_t)cap1 * slots1));
    int tok = prof_begin(stream, PK_FIPPI);
    HIPCHK(launch_fippi(stream, d_jobs_, nt, d_records_, hd, hs, hk, cap1, slots1, d_counter, d_ovf, d_todo));
    prof_end(stream, tok, 0.0);
    std::vector<int> ovf(nj);
    HIPCHK(hipMemcpyAsync(ovf.data(), d_ovf, sizeof(int) * nj, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    hipFree(hd); hipFree(hk); hipFree(hs); hipFree(d_todo);
    std::vector<int> redo;
    for (int j = 0; j < nj; j++) if (ovf[j]) redo.push_back(j);
    if (!redo.empty()) {
      // pass 2: the jobs that outgrew the small heap, heaps large enough for
      // the whole evaluation budget (1e6 points / 66 per split), up to 1024
      // waves (about 20 MB of heap each)
      int cap2 = 32768, slots2 = std::min<int>(1024, (int)redo.size());
      int *d_list = nullptr;
      HIPCHK(hipMalloc(&d_list, sizeof(int) * redo.size()));
      HIPCHK(hipMemcpyAsync(d_list, redo.data(), sizeof(int) * redo.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemsetAsync(d_counter, 0, sizeof(int), stream));
      HIPCHK(hipMemsetAsync(d_ovf, 0, sizeof(int) * nj, stream));
      HIPCHK(hipMalloc(&hd, sizeof(double) * (size_t)cap2 * 76 * slots2));
      HIPCHK(hipMalloc(&hk, sizeof(double) * (size_t)cap2 * slots2));
      HIPCHK(hipMalloc(&hs, sizeof(int) * (size_t)cap2 * slots2));
      tok = prof_begin(stream, PK_FIPPI);
      HIPCHK(launch_fippi(stream, d_jobs_, (int)redo.size(), d_records_, hd, hs, hk, cap2, slots2, d_counter, d_ovf, d_list));
      prof_end(stream, tok, 0.0);
      HIPCHK(hipMemcpyAsync(ovf.data(), d_ovf, sizeof(int) * nj, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      hipFree(hd); hipFree(hk); hipFree(hs); hipFree(d_list);
      for (int j = 0; j < nj; j++) if (ovf[j]) { err_ = "FIPPI cubature exceeded its region budget"; hipFree(d_counter); hipFree(d_ovf); return -4; }
    }
    hipFree(d_counter); hipFree(d_ovf);
    if (prof_enabled()) prof_collect();
    // the new records join the process-wide cache (StoreCache writes them)
    HIPCHK(hipMemcpy(rec.data(), d_records_, sizeof(double) * SCUFFX_QIF_LEN * nj, hipMemcpyDeviceToHost));
    for (int j : todo) fippi_cache_insert(job_keys_[j], rec.data() + (size_t)j * SCUFFX_QIF_LEN);
  }
  fippi_ready_ = true;
  stats_.fippi_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

std::vector<double> Assembler::fippi_records()
{
  std::vector<double> r(jobs_.size() * SCUFFX_QIF_LEN);
  if (!r.empty() && d_records_) hipMemcpy(r.data(), d_records_, sizeof(double) * r.size(), hipMemcpyDeviceToHost);
  return r;
}

static cdouble csqrt2(cdouble z)
{
  cdouble r = std::sqrt(z);
  return r.imag() < 0.0 ? -r : r;
}

// common regions of two surfaces and the relative sign of their
// currents there (SurfaceSurfaceInteractions.cc:62-90)
static int common_regions(const Surface &Sa, const Surface &Sb, int CR[2], double Sg[2])
{
  int n = 0;
  if (Sa.RegionIndices[0] == Sb.RegionIndices[0]) { CR[n] = Sa.RegionIndices[0]; Sg[n] = +1.0; n++; }
  else if (Sa.RegionIndices[0] == Sb.RegionIndices[1]) { CR[n] = Sa.RegionIndices[0]; Sg[n] = -1.0; n++; }
  if (Sa.RegionIndices[1] == Sb.RegionIndices[0]) { CR[n] = Sa.RegionIndices[1]; Sg[n] = -1.0; n++; }
  else if (!Sa.IsPEC && !Sb.IsPEC && Sa.RegionIndices[1] == Sb.RegionIndices[1]) { CR[n] = Sa.RegionIndices[1]; Sg[n] = +1.0; n++; }
  return n;
}

// regions shared by surfaces sa and sb, with wavenumber and the three
// prefactors of the 2x2 block (SurfaceSurfaceInteractions.cc:62-90,
// :162-176)
int Assembler::regions_for(int sa, int sb, const std::vector<cdouble> &eps, const std::vector<cdouble> &mu,
                           cdouble omega, RegionK rk[2])
{
  const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
  int CR[2]; double Sg[2];
  int nreg = common_regions(Sa, Sb, CR, Sg);
  int used = 0;
  for (int r = 0; r < nreg; r++) {
    if (region_only_ >= 0 && CR[r] != region_only_) continue;
    cdouble E = eps[CR[r]], Mu = mu[CR[r]];
    if (E == 0.0) continue;
    cdouble k = csqrt2(E * Mu) * omega, I(0, 1);
    cdouble pf1 = Sg[r] * I * Mu * omega, pf2 = -Sg[r] * I * k, pf3 = -Sg[r] * I * E * omega;
    rk[used].k = mk(k.real(), k.imag());
    rk[used].pf1 = mk(pf1.real(), pf1.imag());
    rk[used].pf2 = mk(pf2.real(), pf2.imag());
    rk[used].pf3 = mk(pf3.real(), pf3.imag());
    used++;
  }
  return used;
}

// Short-wavelength panel pairs (PanelPanelInteractions.cc:426-491): a near
// pair whose larger panel has |k| R > pi leaves the desingularization
// branch -- no common vertex: 78-point product rule; common edge or
// vertex: Taylor-Duffy.  (Self panels stay on Taylor-Duffy with the
// Helmholtz kernel, which the self kernel already does, up to |k| R = 20.)
// The near kernel has already written every record; these jobs overwrite
// the affected ones.
int Assembler::run_sw(int sa, int sb, int bi, const RegionK *rk, int nreg, const RegionK *dreg, int needH1,
                      hipStream_t stream)
{
  const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
  const NearBlock &B = blocks_[bi];
  const int same = (sa == sb);
  std::vector<SWJob> jobs;
  std::vector<char> seen(B.pairs.size(), 0);
  for (int r = 0; r < nreg; r++) {
    const double ak = std::hypot(rk[r].k.r, rk[r].k.i);
    std::fill(seen.begin(), seen.end(), 0);
    for (int pa = 0; pa + 1 < (int)B.rowPtr.size(); pa++)
      for (int e = B.rowPtr[pa]; e < B.rowPtr[pa + 1]; e++) {
        const int id = B.idx[e];
        if (id < 0 || seen[id]) continue;
        seen[id] = 1;
        const int pb = B.col[e];
        const double kR = ak * std::max(Sa.Panels[pa].Radius, Sb.Panels[pb].Radius);
        if (!(kR > M_PI)) continue;
        const int ncv = B.pairs[id].ncv;
        if (ncv == 0) { jobs.push_back(SWJob{id, r, 0, 0}); continue; }
        if (kR > 20.0) {
          err_ = "common-edge/vertex pair with |k| R_panel > 20: the reference's high-k gradient kernel "
                 "has no defined value there (TaylorDuffy.cc:812)";
          return -5;
        }
        for (int q = 0; q < 9; q++) jobs.push_back(SWJob{id, r, q, 1});
      }
    // a self panel past |k| R > 20 (high-k Taylor-Duffy) always comes
    // with common-edge neighbours past 20 on a closed mesh, rejected above
  }
  if (jobs.empty()) return 0;
  if (jobs.size() > cap_swjobs_) {
    if (d_swjobs_) hipFree(d_swjobs_);
    d_swjobs_ = nullptr;
    cap_swjobs_ = 0;
    if (hipMalloc(&d_swjobs_, sizeof(SWJob) * jobs.size()) != hipSuccess) { err_ = "device allocation failed"; return -2; }
    cap_swjobs_ = jobs.size();
  }
  HIPCHK(hipMemcpy(d_swjobs_, jobs.data(), sizeof(SWJob) * jobs.size(), hipMemcpyHostToDevice));
  // one wave per job; 4096 jobs (16 waves per CU) hold 415 MB of value caches
  const int chunk = 4096;
  if (ensure(&d_scratch_, &cap_scratch_, (size_t)chunk * SCUFFX_SW_SLAB)) return -2;
  HIPCHK(launch_sw(stream, d_swjobs_, (int)jobs.size(), B.d_pairs, ds_[sa].V, ds_[sa].PVI, nreg, dreg, needH1,
                   d_HN_, d_HS_, d_scratch_, chunk));
  stats_.nsw += (long)jobs.size();
  return 0;
}

int Assembler::assemble(cdouble omega, void *Mp, long ldm, int kind, hipStream_t stream)
{
  if (!err_.empty()) return -3;
  if (kind != MAT_F64_PACKED && ldm < G_.TotalBFs) { err_ = "leading dimension smaller than the matrix"; return -1; }
  const MatOut M{Mp, ldm, kind};
  int rc = build_fippi(stream);
  if (rc) return rc;
  const int NS = (int)G_.Surfaces.size();
  std::vector<cdouble> eps(G_.Regions.size()), mu(G_.Regions.size());
  for (size_t r = 0; r < G_.Regions.size(); r++) G_.Regions[r].eps_mu(omega, &eps[r], &mu[r]);
  hipEvent_t ev[6];
  for (auto &e : ev) hipEventCreate(&e);
  double t_self = 0, t_far = 0, t_near = 0, t_stamp = 0;
  int bi = 0;
  stats_.nfar = stats_.nnear = stats_.nself = stats_.nsw = 0;
  for (int sa = 0; sa < NS; sa++)
    for (int sb = sa; sb < NS; sb++, bi++) {
      const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
      const DevSurface &Da = ds_[sa], &Db = ds_[sb];
      NearBlock &B = blocks_[bi];
      const int same = (sa == sb);
      if (same && mates_[sa] >= 0) {
        // identical to an earlier surface: its diagonal block is copied
        // (AssembleBEMMatrix.cc:733-740); the lower triangle comes from the
        // final mirror as for every diagonal block
        const long om = G_.BFIndexOffset[mates_[sa]], os = G_.BFIndexOffset[sa], n = Sa.NumBFs;
        HIPCHK(launch_copy_block(stream, M, om, os, n));
        continue;
      }
      RegionK rk[2];
      int nreg = regions_for(sa, sb, eps, mu, omega, rk);
      if (nreg < 0) {
        for (auto &e : ev) hipEventDestroy(e);
        return nreg;
      }
      const int needH1 = !(Sa.IsPEC && Sb.IsPEC);
      const int stride = needH1 ? 4 : 2;
      HIPCHK(hipMemcpyAsync(d_regs_ + 2 * (bi % 2), rk, sizeof(RegionK) * std::max(nreg, 1), hipMemcpyHostToDevice, stream));
      RegionK *dreg = d_regs_ + 2 * (bi % 2);
      if (ensure(&d_HN_, &cap_HN_, (size_t)B.pairs.size() * nreg * 9 * stride)) return -2;
      if (same) {
        if (ensure(&d_HS_, &cap_HS_, (size_t)Da.nP * nreg * 18)) return -2;
        int chunk = 16384;
        if (ensure(&d_scratch_, &cap_scratch_, (size_t)chunk * 1025 * 4)) return -2;
        hipEventRecord(ev[0], stream);
        int tok = prof_begin(stream, PK_SELF);
        if (nreg > 0) HIPCHK(launch_self(stream, Da.V, Da.PVI, Da.nP, nreg, dreg, d_HS_, d_scratch_, chunk));
        prof_end(stream, tok, 0.0);
        hipEventRecord(ev[1], stream);
        stats_.nself += (long)Da.nP * 9 * nreg;
      } else {
        hipEventRecord(ev[0], stream);
        hipEventRecord(ev[1], stream);
      }
      const double wpair = needH1 ? 3780.0 : 2160.0;
      const long nfar_blk = B.bands.uniqueFar;
      if (nreg > 0 && !B.pairs.empty()) {
        int tok = prof_begin(stream, PK_NEAR);
        HIPCHK(launch_near(stream, B.d_pairs, (int)B.pairs.size(), d_records_, nreg, dreg, needH1, d_HN_));
        prof_end(stream, tok, wpair * (double)B.pairs.size() * nreg);
      }
      if (nreg > 0) {
        int rc2 = run_sw(sa, sb, bi, rk, nreg, dreg, needH1, stream);
        if (rc2) {
          for (auto &e : ev) hipEventDestroy(e);
          return rc2;
        }
      }
      hipEventRecord(ev[2], stream);
      HIPCHK(hipEventSynchronize(ev[2]));
      float ms;
      hipEventElapsedTime(&ms, ev[0], ev[1]); t_self += ms;
      hipEventElapsedTime(&ms, ev[1], ev[2]); t_near += ms;
      // far records and matrix entries, one column band at a time
      int rc3 = far_and_stamp(B, Da, Db, Sa, Sb, nreg, dreg, needH1, M, G_.BFIndexOffset[sa], G_.BFIndexOffset[sb],
                              stream, &t_far, &t_stamp);
      if (rc3) {
        for (auto &e : ev) hipEventDestroy(e);
        return rc3;
      }
      stats_.nfar += nfar_blk;
      stats_.nnear += (long)B.pairs.size();
    }
  hipEventRecord(ev[0], stream);
  HIPCHK(launch_mirror(stream, M, G_.TotalBFs));
  hipEventRecord(ev[1], stream);
  HIPCHK(hipEventSynchronize(ev[1]));
  float ms;
  hipEventElapsedTime(&ms, ev[0], ev[1]);
  stats_.mirror_ms = ms;
  stats_.self_ms = t_self; stats_.far_ms = t_far; stats_.near_ms = t_near; stats_.stamp_ms = t_stamp;
  for (auto &e : ev) hipEventDestroy(e);
  if (prof_enabled()) prof_collect();
  return 0;
}

int Assembler::prepare_graph(const std::vector<cdouble> &omegas, hipStream_t stream)
{
  if (!err_.empty()) return -3;
  int rc = build_fippi(stream);
  if (rc) return rc;
  const int NS = (int)G_.Surfaces.size(), nf = (int)omegas.size();
  const int nblk = NS * (NS + 1) / 2;
  std::vector<RegionK> tab((size_t)std::max(nf, 1) * nblk * 2);
  nreg_tab_.assign((size_t)std::max(nf, 1) * nblk, 0);
  size_t need_hn = 1, need_hs = 1, need_hb = 1;
  for (int f = 0; f < nf; f++) {
    std::vector<cdouble> eps(G_.Regions.size()), mu(G_.Regions.size());
    for (size_t r = 0; r < G_.Regions.size(); r++) G_.Regions[r].eps_mu(omegas[f], &eps[r], &mu[r]);
    int bi = 0;
    for (int sa = 0; sa < NS; sa++)
      for (int sb = sa; sb < NS; sb++, bi++) {
        RegionK rk[2];
        const int nreg = regions_for(sa, sb, eps, mu, omegas[f], rk);
        if (nreg < 0) return nreg;
        nreg_tab_[(size_t)f * nblk + bi] = nreg;
        for (int r = 0; r < nreg; r++) tab[((size_t)f * nblk + bi) * 2 + r] = rk[r];
        // the short-wavelength kernels build their job lists on the host
        const NearBlock &B = blocks_[bi];
        for (int r = 0; r < nreg; r++) {
          const double ak = std::hypot(rk[r].k.r, rk[r].k.i);
          const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
          for (int pa = 0; pa + 1 < (int)B.rowPtr.size(); pa++)
            for (int e = B.rowPtr[pa]; e < B.rowPtr[pa + 1]; e++)
              if (B.idx[e] >= 0 && ak * std::max(Sa.Panels[pa].Radius, Sb.Panels[B.col[e]].Radius) > M_PI) {
                err_ = "a sweep frequency needs the short-wavelength kernels (|k| R_panel > pi): not capturable";
                return -5;
              }
        }
        const int needH1 = !(G_.Surfaces[sa].IsPEC && G_.Surfaces[sb].IsPEC);
        const size_t stride = needH1 ? 4 : 2;
        need_hn = std::max(need_hn, B.pairs.size() * std::max(nreg, 1) * 9 * stride);
        if (sa == sb) need_hs = std::max(need_hs, (size_t)ds_[sa].nP * std::max(nreg, 1) * 18);
        need_hb = std::max(need_hb, (size_t)std::max(B.bands.maxSlots, 1) * B.bands.rowsCap * std::max(nreg, 1) * 9 * stride);
      }
  }
  if (ensure(&d_HN_, &cap_HN_, need_hn) || ensure(&d_HS_, &cap_HS_, need_hs) || ensure(&d_HP_, &cap_HP_, need_hb) ||
      ensure(&d_scratch_, &cap_scratch_, (size_t)16384 * 1025 * 4))
    return -2;
  const size_t bytes = sizeof(RegionK) * tab.size();
  if (bytes > cap_regtab_) {
    alloc_gen_++;
    if (d_regtab_) hipFree(d_regtab_);
    d_regtab_ = nullptr;
    cap_regtab_ = 0;
    if (hipMalloc(&d_regtab_, bytes) != hipSuccess) { err_ = "device allocation failed"; return -2; }
    cap_regtab_ = bytes;
  }
  HIPCHK(hipMemcpy(d_regtab_, tab.data(), bytes, hipMemcpyHostToDevice));
  return 0;
}

int Assembler::assemble_graph(int f, void *Mp, long ldm, hipStream_t stream)
{
  if (!err_.empty()) return -3;
  if (!d_regtab_ || ldm < G_.TotalBFs) { err_ = "sweep not prepared"; return -1; }
  const MatOut M{Mp, ldm, MAT_C128};
  const int NS = (int)G_.Surfaces.size(), nblk = NS * (NS + 1) / 2;
  int bi = 0;
  for (int sa = 0; sa < NS; sa++)
    for (int sb = sa; sb < NS; sb++, bi++) {
      const Surface &Sa = G_.Surfaces[sa], &Sb = G_.Surfaces[sb];
      const DevSurface &Da = ds_[sa], &Db = ds_[sb];
      NearBlock &B = blocks_[bi];
      const int same = (sa == sb);
      if (same && mates_[sa] >= 0) {
        const long om = G_.BFIndexOffset[mates_[sa]], os = G_.BFIndexOffset[sa], n = Sa.NumBFs;
        HIPCHK(launch_copy_block(stream, M, om, os, n));
        continue;
      }
      const int nreg = nreg_tab_[(size_t)f * nblk + bi];
      const RegionK *dreg = d_regtab_ + ((size_t)f * nblk + bi) * 2;
      const int needH1 = !(Sa.IsPEC && Sb.IsPEC);
      if (same && nreg > 0)
        HIPCHK(launch_self(stream, Da.V, Da.PVI, Da.nP, nreg, dreg, d_HS_, d_scratch_, 16384));
      if (nreg > 0 && !B.pairs.empty())
        HIPCHK(launch_near(stream, B.d_pairs, (int)B.pairs.size(), d_records_, nreg, dreg, needH1, d_HN_));
      int rc = far_and_stamp(B, Da, Db, Sa, Sb, nreg, dreg, needH1, M, G_.BFIndexOffset[sa], G_.BFIndexOffset[sb],
                             stream, nullptr, nullptr);
      if (rc) return rc;
    }
  HIPCHK(launch_mirror(stream, M, G_.TotalBFs));
  return 0;
}

__global__ void k_transpose_into(const double2 *T, long ldt, int rows, int cols, double2 *B, long ldb)
{
  __shared__ double2 tile[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;   // tile of T: rows bx.., cols by..
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    int r = bx + tx, c = by + yy;
    if (r < rows && c < cols) tile[yy][tx] = T[r + (long)c * ldt];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int r = by + tx, c = bx + yy;              // B(r, c) = T(c, r)
    if (r < cols && c < rows) B[r + (long)c * ldb] = tile[tx][yy];
  }
}

// AssembleBEMMatrixBlock (libscuff.h:500-505, AssembleBEMMatrix.cc:349-391):
// the block B = M(rows of sa, columns of sb) and, for two distinct
// surfaces, if GB is given, its derivatives with respect to a rigid
// displacement of surface sa (three matrices gstride elements apart).
// sa == sb gives the diagonal block (the T block of the Casimir caller,
// CasimirIntegrand.cc:276-291) without derivatives.
int Assembler::assemble_block(int sa, int sb, cdouble omega, double2 *B, long ldb, double2 *GB, long ldg,
                              long gstride, int ntorque, const double *gamma, double2 *TB, hipStream_t stream)
{
  if (ntorque < 0 || ntorque > 3 || (ntorque > 0 && (!gamma || !TB || !GB))) {
    err_ = "torque derivatives need 1 to 3 axes, their generators and the gradient output";
    return -1;
  }
  if (!err_.empty()) return -3;
  const int NS = (int)G_.Surfaces.size();
  if (sa < 0 || sb < 0 || sa >= NS || sb >= NS) { err_ = "surface index out of range"; return -1; }
  const int same = (sa == sb);
  if (same && GB) { err_ = "derivatives need two distinct surfaces"; return -1; }
  const Surface &SA = G_.Surfaces[sa], &SB = G_.Surfaces[sb];
  if (ldb < SA.NumBFs || (GB && ldg < SA.NumBFs) || (GB && gstride < ldg * (long)SB.NumBFs)) {
    err_ = "leading dimension smaller than the block";
    return -1;
  }
  int rc = build_fippi(stream);
  if (rc) return rc;
  std::vector<cdouble> eps(G_.Regions.size()), mu(G_.Regions.size());
  for (size_t r = 0; r < G_.Regions.size(); r++) G_.Regions[r].eps_mu(omega, &eps[r], &mu[r]);
  const int lo = std::min(sa, sb), hi = std::max(sa, sb);
  int bi = 0;
  for (int i = 0; i < lo; i++) bi += NS - i;
  bi += hi - lo;
  const Surface &Sa = G_.Surfaces[lo], &Sb = G_.Surfaces[hi];
  const DevSurface &Da = ds_[lo], &Db = ds_[hi];
  NearBlock &NB = blocks_[bi];
  RegionK rk[2];
  int nreg = regions_for(lo, hi, eps, mu, omega, rk);
  if (nreg < 0) return nreg;
  const int needH1 = !(Sa.IsPEC && Sb.IsPEC);
  const int stride = needH1 ? 4 : 2;
  HIPCHK(hipMemcpyAsync(d_regs_, rk, sizeof(RegionK) * std::max(nreg, 1), hipMemcpyHostToDevice, stream));
  if (ensure(&d_HN_, &cap_HN_, (size_t)NB.pairs.size() * nreg * 9 * stride)) return -2;
  if (same && nreg > 0) {
    const int chunk = 16384;
    if (ensure(&d_HS_, &cap_HS_, (size_t)Da.nP * nreg * 18)) return -2;
    if (ensure(&d_scratch_, &cap_scratch_, (size_t)chunk * 1025 * 4)) return -2;
    HIPCHK(launch_self(stream, Da.V, Da.PVI, Da.nP, nreg, d_regs_, d_HS_, d_scratch_, chunk));
  }
  // the block in (lo, hi) orientation, directly into B or into a
  // staging copy that is transposed into B
  double2 *target = B;
  long ldt = ldb;
  double *stage = nullptr;
  if (sa > sb) {
    if (ensure(&d_stage_, &cap_stage_, (size_t)2 * Sa.NumBFs * Sb.NumBFs)) return -2;
    stage = d_stage_;
    target = (double2 *)stage;
    ldt = Sa.NumBFs;
  }
  if (nreg > 0) {
    if (!NB.pairs.empty())
      HIPCHK(launch_near(stream, NB.d_pairs, (int)NB.pairs.size(), d_records_, nreg, d_regs_, needH1, d_HN_));
    int rc2 = run_sw(lo, hi, bi, rk, nreg, d_regs_, needH1, stream);
    if (rc2) return rc2;
  }
  {
    int rc3 = far_and_stamp(NB, Da, Db, Sa, Sb, nreg, d_regs_, needH1, MatOut{target, ldt, MAT_C128}, 0, 0, stream,
                            nullptr, nullptr);
    if (rc3) return rc3;
  }
  if (same) HIPCHK(launch_mirror(stream, MatOut{B, ldb, MAT_C128}, Sa.NumBFs));
  if (stage) {
    dim3 grid((Sa.NumBFs + 31) / 32, (Sb.NumBFs + 31) / 32);
    hipLaunchKernelGGL(k_transpose_into, grid, dim3(256), 0, stream, (const double2 *)stage, ldt, Sa.NumBFs,
                       Sb.NumBFs, B, ldb);
    HIPCHK(hipGetLastError());
  }
  if (GB) {
    // gradient in the caller's orientation; regions and prefactors of (sa, sb)
    RegionK gk[2];
    int ng = regions_for(sa, sb, eps, mu, omega, gk);
    if (ng < 0) return ng;
    HIPCHK(hipMemcpyAsync(d_regs_ + 2, gk, sizeof(RegionK) * std::max(ng, 1), hipMemcpyHostToDevice, stream));
    const DevSurface &DA = ds_[sa], &DB = ds_[sb];
    int tok = prof_begin(stream, PK_GRAD);
    GammaSet GS;
    std::memset(&GS, 0, sizeof(GS));
    GS.n = ntorque;
    for (int i = 0; i < 9 * ntorque; i++) GS.g[i] = gamma[i];
    HIPCHK(launch_grad(stream, DA.nE, DA.EI, DA.EL, DA.V, DA.PVI, DA.PC, DB.nE, DB.EI, DB.EL, DB.V, DB.PVI, DB.PC,
                       SA.IsPEC, SB.IsPEC, ng, d_regs_ + 2, GB, ldg, gstride, GS, TB));
    prof_end(stream, tok, 0.0);
  }
  HIPCHK(hipStreamSynchronize(stream));
  if (prof_enabled()) prof_collect();
  return 0;
}

}  // namespace scuffx
