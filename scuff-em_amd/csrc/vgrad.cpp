// vgrad.cpp -- host orchestration of the vertex-position derivative of the
// BEM matrix contracted with weights C (kernels in kern_vgrad.hip):
//   grad(theta) = sum_ij C_ij dM_ij / dtheta
// for every vertex coordinate theta of every surface.  Panel pairs are
// split as the assembly splits them (AssessPanelPair.cc:52-172; far when
// the relative centroid distance exceeds 4):
//   far pairs           k_vg_far, analytic, degree-4 rule (the assembly's)
//   near, no common     k_vg_nearsep, analytic, 78-point rule per quarter
//   common edge/vertex  central differences of Taylor-Duffy integrals at
//                       displaced vertices (k_sw), contracted by k_vg_fd
//   a panel with itself central differences of k_self, likewise
// plus the edge-length term from the assembled matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "assemble.h"
#include "vgrad.h"
#include "ws.h"

namespace scuffx {

namespace {

// device buffers of one call, freed on return
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch()
  {
    if (hip_down()) return;
    for (void *p : ptrs) hipFree(p);
  }
  template <typename T>
  T *alloc(size_t n)
  {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T *)p;
  }
  template <typename T>
  T *put(const std::vector<T> &h)
  {
    T *d = alloc<T>(h.size());
    if (d && !h.empty() && hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    return d;
  }
};

// relative step of the central differences, in units of the larger
// panel radius of the pair
constexpr double kFDStep = 1.0e-4;

}  // namespace

int Assembler::vg_tables()
{
  if (!vgs_.empty()) return 0;
  const int NS = (int)G_.Surfaces.size();
  std::vector<VGSurf> vs(NS);
  long gofs = 0;
  for (int s = 0; s < NS; s++) {
    const Surface &S = G_.Surfaces[s];
    const int nP = S.NumPanels(), nE = S.NumEdges();
    std::vector<int> pbf(3 * (size_t)nP, -1), ev(4 * (size_t)std::max(nE, 1), -1);
    std::vector<double> psg(3 * (size_t)nP, 0.0);
    for (int e = 0; e < nE; e++) {
      const Edge &E = S.Edges[e];
      pbf[3 * E.iPPanel + E.PIndex] = e;
      psg[3 * E.iPPanel + E.PIndex] = 1.0;
      if (E.iMPanel >= 0) {
        pbf[3 * E.iMPanel + E.MIndex] = e;
        psg[3 * E.iMPanel + E.MIndex] = -1.0;
      }
      ev[4 * e] = E.iV1; ev[4 * e + 1] = E.iV2; ev[4 * e + 2] = E.iQP; ev[4 * e + 3] = E.iQM;
    }
    int *dpbf = nullptr, *dev = nullptr;
    double *dpsg = nullptr;
    if (hipMalloc(&dpbf, sizeof(int) * pbf.size()) != hipSuccess || hipMalloc(&dpsg, sizeof(double) * psg.size()) != hipSuccess ||
        hipMalloc(&dev, sizeof(int) * ev.size()) != hipSuccess) {
      hipFree(dpbf); hipFree(dpsg); hipFree(dev);
      return -2;
    }
    vg_alloc_.push_back(dpbf); vg_alloc_.push_back(dpsg); vg_alloc_.push_back(dev);
    if (hipMemcpy(dpbf, pbf.data(), sizeof(int) * pbf.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dpsg, psg.data(), sizeof(double) * psg.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dev, ev.data(), sizeof(int) * ev.size(), hipMemcpyHostToDevice) != hipSuccess)
      return -3;
    VGSurf &A = vs[s];
    A.V = ds_[s].V; A.PVI = ds_[s].PVI; A.EL = ds_[s].EL;
    A.PBF = dpbf; A.PSG = dpsg;
    A.nP = nP; A.nV = S.NumVertices(); A.off = G_.BFIndexOffset[s]; A.pec = S.IsPEC ? 1 : 0;
    A.gofs = gofs;
    gofs += 3L * A.nV;
    vg_ev_.push_back(dev);
  }
  vg_slots_ = gofs;
  vgs_ = vs;
  return 0;
}

int Assembler::vertex_gradient(cdouble omega, const double2 *M, long ldm, const VGW &w, double2 *grad,
                               hipStream_t stream, std::string *msg)
{
  auto bad = [&](int rc, const char *m) { if (msg) *msg = m; return rc; };
  if (!err_.empty()) return bad(-3, "assembler in error");
  if (vg_tables()) return bad(-2, "device allocation failed");
  const int NS = (int)G_.Surfaces.size();
  const int nblk = NS * (NS + 1) / 2;
  auto bidx = [&](int a, int b) {
    if (a > b) std::swap(a, b);
    return a * NS - a * (a - 1) / 2 + (b - a);
  };
  std::vector<cdouble> eps(G_.Regions.size()), mu(G_.Regions.size());
  for (size_t r = 0; r < G_.Regions.size(); r++) G_.Regions[r].eps_mu(omega, &eps[r], &mu[r]);
  std::vector<RegionK> rk(2 * (size_t)nblk);
  std::vector<int> nreg(nblk, 0);
  for (int sa = 0; sa < NS; sa++)
    for (int sb = sa; sb < NS; sb++) {
      const int b = bidx(sa, sb);
      nreg[b] = regions_for(sa, sb, eps, mu, omega, &rk[2 * b]);
      if (nreg[b] < 0) return bad(-1, "region constants failed");
    }
  Scratch sc;
  RegionK *drk = sc.put(rk);
  VGSurf *dvs = sc.put(vgs_);
  if (!drk || !dvs) return bad(-2, "device allocation failed");
  if (hipMemsetAsync(grad, 0, sizeof(double2) * vg_slots_, stream) != hipSuccess) return bad(-3, "HIP error");

  // ---- near lists per ordered surface pair; far and separated-near kernels
  struct TouchPair { int s, p, q, ncv, ia[3], ib[3]; };
  std::vector<TouchPair> touch;
  for (int s = 0; s < NS; s++)
    for (int t = 0; t < NS; t++) {
      const Surface &Ss = G_.Surfaces[s], &St = G_.Surfaces[t];
      const int b = bidx(s, t);
      const NearBlock &B = blocks_[b];
      std::vector<int> rowPtr, col;
      if (s <= t) {
        rowPtr = B.rowPtr;
        col = B.col;
      } else {   // rows of s: transpose of block (t, s)
        const int nPs = Ss.NumPanels();
        std::vector<std::vector<int>> rows(nPs);
        for (int pa = 0; pa + 1 < (int)B.rowPtr.size(); pa++)
          for (int e = B.rowPtr[pa]; e < B.rowPtr[pa + 1]; e++) rows[B.col[e]].push_back(pa);
        rowPtr.assign(nPs + 1, 0);
        for (int p = 0; p < nPs; p++) {
          std::sort(rows[p].begin(), rows[p].end());
          rowPtr[p + 1] = rowPtr[p] + (int)rows[p].size();
          col.insert(col.end(), rows[p].begin(), rows[p].end());
        }
      }
      std::vector<int> sep;
      for (int p = 0; p + 1 < (int)rowPtr.size(); p++)
        for (int e = rowPtr[p]; e < rowPtr[p + 1]; e++) {
          const int q = col[e];
          int ia[3], ib[3];
          double rRel;
          const int ncv = assess_pair(Ss, p, St, q, &rRel, ia, ib);
          if (ncv == 0) { sep.push_back(p); sep.push_back(q); continue; }
          if (s != t) return bad(-4, "touching surfaces: the vertex gradient needs each shared vertex on one surface");
          if (p < q) {
            TouchPair T{s, p, q, ncv, {ia[0], ia[1], ia[2]}, {ib[0], ib[1], ib[2]}};
            touch.push_back(T);
          }
        }
      if (nreg[b] == 0) continue;
      const int needH1 = !(Ss.IsPEC && St.IsPEC);
      int *drp = sc.put(rowPtr), *dcol = sc.put(col), *dsep = sc.put(sep);
      if (!drp || !dcol || !dsep) return bad(-2, "device allocation failed");
      if (launch_vg_far(stream, vgs_[s], vgs_[t], s == t, drp, dcol, nreg[b], drk + 2 * b, needH1, w, grad) != hipSuccess ||
          launch_vg_nearsep(stream, vgs_[s], vgs_[t], dsep, (int)sep.size() / 2, nreg[b], drk + 2 * b, needH1, w,
                            grad) != hipSuccess)
        return bad(-3, "vertex-gradient kernel failed");
    }

  // ---- touching pairs: displaced Taylor-Duffy evaluations
  bool anyH1 = false;
  for (const Surface &S : G_.Surfaces) anyH1 = anyH1 || !S.IsPEC;
  const int stride = anyH1 ? 4 : 2;
  std::vector<VGFDItem> items;
  // per surface: its NearPair list and jobs (one k_sw launch each)
  struct SurfJobs { std::vector<NearPair> np; long base = 0; };
  std::vector<SurfJobs> sj(NS);
  long hnTotal = 0;
  for (int s = 0; s < NS; s++) {
    const Surface &S = G_.Surfaces[s];
    const int b = bidx(s, s), nr = nreg[b];
    sj[s].base = hnTotal;
    if (nr == 0) continue;
    for (const TouchPair &T : touch) {
      if (T.s != s) continue;
      const Panel &Pp = S.Panels[T.p], &Pq = S.Panels[T.q];
      const double h = kFDStep * std::max(Pp.Radius, Pq.Radius);
      int verts[6], nv = 0;
      for (int i = 0; i < 3; i++) verts[nv++] = Pp.VI[i];
      for (int i = 0; i < 3; i++)
        if (Pq.VI[i] != Pp.VI[0] && Pq.VI[i] != Pp.VI[1] && Pq.VI[i] != Pp.VI[2]) verts[nv++] = Pq.VI[i];
      for (int m = 0; m < nv; m++)
        for (int c = 0; c < 3; c++) {
          VGFDItem it;
          it.sa = s; it.pa = T.p; it.sb = s; it.pb = T.q; it.blk = b; it.nreg = nr; it.self = 0;
          it.stride = stride; it.hofs = hnTotal + (long)sj[s].np.size() * nr * 9 * stride;
          it.slot = vgs_[s].gofs + 3L * verts[m] + c; it.h = h;
          items.push_back(it);
          for (int sgn = 0; sgn < 2; sgn++) {
            const double dx = sgn == 0 ? h : -h;
            auto X = [&](int vi, int cc) { return S.V[3 * vi + cc] + (vi == verts[m] && cc == c ? dx : 0.0); };
            NearPair P;
            std::memset(&P, 0, sizeof(P));
            for (int i = 0; i < 3; i++)
              for (int cc = 0; cc < 3; cc++) {
                P.Va[3 * i + cc] = X(Pp.VI[T.ia[i]], cc);
                P.Vb[3 * i + cc] = X(Pq.VI[T.ib[i]], cc);
                P.Qa[3 * i + cc] = X(Pp.VI[i], cc);
                P.Qb[3 * i + cc] = X(Pq.VI[i], cc);
              }
            P.ncv = T.ncv;
            P.rec = -1;
            sj[s].np.push_back(P);
          }
        }
    }
    hnTotal += (long)sj[s].np.size() * nr * 9 * stride;
  }
  // ---- panels with themselves: displaced self integrals
  std::vector<double> synV;
  std::vector<int> synPVI;
  std::vector<long> synBase(NS, 0), synCount(NS, 0);
  long hsTotal = 0;
  for (int s = 0; s < NS; s++) {
    const Surface &S = G_.Surfaces[s];
    const int b = bidx(s, s), nr = nreg[b];
    synBase[s] = (long)synPVI.size() / 3;
    if (nr == 0) continue;
    for (int p = 0; p < S.NumPanels(); p++) {
      const Panel &P = S.Panels[p];
      const double h = kFDStep * P.Radius;
      for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
          VGFDItem it;
          const long j = (long)synPVI.size() / 3;   // the +h synthetic panel
          it.sa = s; it.pa = p; it.sb = s; it.pb = p; it.blk = b; it.nreg = nr; it.self = 1; it.stride = 2;
          it.hofs = hsTotal + (j - synBase[s]) * nr * 18;
          it.slot = vgs_[s].gofs + 3L * P.VI[k] + c; it.h = h;
          items.push_back(it);
          for (int sgn = 0; sgn < 2; sgn++) {
            const long jj = (long)synPVI.size() / 3;
            for (int i = 0; i < 3; i++) {
              for (int cc = 0; cc < 3; cc++)
                synV.push_back(S.V[3 * P.VI[i] + cc] + (i == k && cc == c ? (sgn == 0 ? h : -h) : 0.0));
              synPVI.push_back((int)(3 * jj + i));
            }
          }
        }
    }
    synCount[s] = (long)synPVI.size() / 3 - synBase[s];
    hsTotal += synCount[s] * nr * 18;
  }
  double *dHN = sc.alloc<double>(hnTotal), *dHS = sc.alloc<double>(hsTotal);
  if (!dHN || !dHS) return bad(-2, "device allocation failed");
  {
    // Taylor-Duffy jobs, one launch per surface
    const int chunk = 4096;
    double *scr = nullptr;
    bool needSW = false;
    for (int s = 0; s < NS; s++) needSW = needSW || !sj[s].np.empty();
    if (needSW && !(scr = sc.alloc<double>((size_t)chunk * SCUFFX_SW_SLAB))) return bad(-2, "device allocation failed");
    for (int s = 0; s < NS; s++) {
      if (sj[s].np.empty()) continue;
      const int b = bidx(s, s), nr = nreg[b];
      std::vector<SWJob> jobs;
      jobs.reserve(sj[s].np.size() * nr * 9);
      for (int n = 0; n < (int)sj[s].np.size(); n++)
        for (int r = 0; r < nr; r++)
          for (int q = 0; q < 9; q++) jobs.push_back(SWJob{n, r, q, 1});
      NearPair *dnp = sc.put(sj[s].np);
      SWJob *djobs = sc.put(jobs);
      if (!dnp || !djobs) return bad(-2, "device allocation failed");
      if (launch_sw(stream, djobs, (int)jobs.size(), dnp, nullptr, nullptr, nr, drk + 2 * b, stride == 4,
                    dHN + sj[s].base, nullptr, scr, chunk) != hipSuccess)
        return bad(-3, "Taylor-Duffy kernel failed");
    }
  }
  if (!synPVI.empty()) {
    const int chunk = 8192;
    double *scr = sc.alloc<double>((size_t)chunk * 1025 * 4), *dV = sc.put(synV);
    int *dP = sc.put(synPVI);
    if (!scr || !dV || !dP) return bad(-2, "device allocation failed");
    for (int s = 0; s < NS; s++) {
      if (synCount[s] == 0) continue;
      const int b = bidx(s, s), nr = nreg[b];
      long hs0 = 0;
      for (int t = 0; t < s; t++) hs0 += synCount[t] * nreg[bidx(t, t)] * 18;
      if (launch_self(stream, dV, dP + 3 * synBase[s], (int)synCount[s], nr, drk + 2 * b, dHS + hs0, scr, chunk) !=
          hipSuccess)
        return bad(-3, "self-term kernel failed");
    }
  }
  if (!items.empty()) {
    VGFDItem *dit = sc.put(items);
    if (!dit) return bad(-2, "device allocation failed");
    if (launch_vg_fd(stream, dvs, dit, (int)items.size(), drk, dHN, dHS, w, grad) != hipSuccess)
      return bad(-3, "difference kernel failed");
  }
  // ---- edge lengths: sum_i rho_i d ln L_i, rho_i = sum_j C_ij M_ij + C_ji M_ji
  const int N = G_.TotalBFs;
  double2 *rho = sc.alloc<double2>(N);
  if (!rho) return bad(-2, "device allocation failed");
  if (hipMemsetAsync(rho, 0, sizeof(double2) * N, stream) != hipSuccess ||
      launch_vg_rho(stream, M, ldm, N, w, rho) != hipSuccess)
    return bad(-3, "edge-length kernel failed");
  for (int s = 0; s < NS; s++)
    if (launch_vg_len(stream, vgs_[s], vg_ev_[s], G_.Surfaces[s].NumEdges(), rho, grad) != hipSuccess)
      return bad(-3, "edge-length kernel failed");
  // the scratch buffers are freed on return: finish first
  if (hipStreamSynchronize(stream) != hipSuccess) return bad(-3, "vertex-gradient kernels failed on the device");
  return 0;
}

}  // namespace scuffx
