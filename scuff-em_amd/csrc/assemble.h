// assemble.h -- host orchestration of the BEM matrix assembly on one GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bem.h"
#include "fippi_cache.h"
#include "geometry.h"
#include "vgrad.h"

namespace scuffx {

struct DevSurface {
  double *V = nullptr;
  int *PVI = nullptr;
  double *PC = nullptr;
  int *EI = nullptr;
  double *EL = nullptr;
  int nV = 0, nP = 0, nE = 0;
};

// Column bands of one surface-pair block (frequency independent).  Band k
// holds the column edges whose higher panel index lies in a range of `ch`
// panels, the sorted list of their panels (the band's slots) and, per
// column edge, the slots of its two panels.  Its far records pair every
// row panel pa < rowLimit[k] with every slot; its stamp covers row edges
// < rowEdges[k] (the upper triangle of a diagonal block ends at the
// band's last column).  One band's records take at most
// maxSlots x rowsCap records, instead of the whole block's panel pairs.
struct FarBands {
  int ch = 0, maxSlots = 0, rowsCap = 0;
  std::vector<int> colStart, cols, colSlots;   // colSlots: 2 per column edge
  std::vector<int> slotStart, slots;
  std::vector<int> rowLimit, rowEdges;
  std::vector<long> farPairs;                  // far pairs computed per band
  long uniqueFar = 0;                          // far panel pairs of the block, each once
  int *d_cols = nullptr, *d_colSlots = nullptr, *d_slots = nullptr;
  int count() const { return (int)rowLimit.size(); }
};

// near-pair bookkeeping of one surface-pair block
struct NearBlock {
  int sa = 0, sb = 0;
  std::vector<int> rowPtr, col, idx;   // CSR over row panels, sorted columns
  std::vector<NearPair> pairs;         // first-use order
  int *d_rowPtr = nullptr, *d_col = nullptr, *d_idx = nullptr;
  NearPair *d_pairs = nullptr;
  FarBands bands;
};

struct AsmStats {
  double fippi_ms = 0, self_ms = 0, far_ms = 0, near_ms = 0, stamp_ms = 0, mirror_ms = 0;
  long nfar = 0, nnear = 0, nself = 0, nfippi = 0, nsw = 0, nfippi_cached = 0;
};

class Assembler {
 public:
  explicit Assembler(const Geometry &G);
  ~Assembler();
  // fills the full N x N matrix in device memory: kind MAT_C128 (column-
  // major complex128), MAT_F64 (real parts, column-major) or
  // MAT_F64_PACKED (real parts, packed upper triangle; ldm unused)
  int assemble(cdouble omega, void *M, long ldm, int kind, hipStream_t stream);
  int build_fippi(hipStream_t stream);
  // Capturable assembly of a frequency sweep.  prepare_graph (outside any
  // stream capture) builds the FIPPI table, checks that no frequency needs
  // the short-wavelength kernels, sizes every buffer for the largest block
  // and uploads each frequency's region constants; assemble_graph(f, ...)
  // then assembles frequency f with no host synchronization, allocation or
  // host-to-device copy (complex128, full matrix).
  int prepare_graph(const std::vector<cdouble> &omegas, hipStream_t stream);
  int assemble_graph(int f, void *M, long ldm, hipStream_t stream);
  // block of surfaces sa (rows) and sb (columns); for sa != sb optionally
  // its derivatives with respect to a rigid motion of surface sa
  int assemble_block(int sa, int sb, cdouble omega, double2 *B, long ldb, double2 *GB, long ldg, long gstride,
                     int ntorque, const double *gamma, double2 *TB, hipStream_t stream);
  // r >= 0: later assemblies keep only the contributions of region r
  // (its terms of every common-region sum); -1: every region
  void set_region_filter(int r) { region_only_ = r; }
  // 32: far panel pairs integrated in single precision by later
  // assemblies (the mixed form's option); 64: double (default)
  void set_far_precision(int bits) { far_bits_ = bits == 32 ? 32 : 64; }
  int far_precision() const { return far_bits_; }
  const AsmStats &stats() const { return stats_; }
  // unique per Assembler object of the process (device pointers recorded
  // in a graph belong to one assembler)
  unsigned long long serial() const { return serial_; }
  // changes whenever a device buffer of the assembler is reallocated
  unsigned long long alloc_generation() const { return alloc_gen_; }
  const std::vector<int> &nreg_table() const { return nreg_tab_; }
  const DevSurface &dev_surface(int s) const { return ds_[s]; }
  // sum_ij C_ij dM_ij / dtheta for every vertex coordinate theta of every
  // surface (grad: complex, slot gofs(s) + 3 v + c, vg_slots() of them);
  // M: the matrix assembled at omega, for the edge-length term; *msg says
  // why on failure (vgrad.cpp)
  int vertex_gradient(cdouble omega, const double2 *M, long ldm, const VGW &w, double2 *grad, hipStream_t stream,
                      std::string *msg);
  // per-surface views for the vertex-gradient kernels, built on first use
  int vg_tables();
  const VGSurf &vg_surface(int s) const { return vgs_[s]; }
  const int *vg_edges(int s) const { return vg_ev_[s]; }   // iV1, iV2, iQP, iQM per edge
  long vg_slots() const { return vg_slots_; }
  std::string error() const { return err_; }
  int num_fippi() const { return (int)jobs_.size(); }
  // host copy of the FIPPI records (36 doubles each) for checks
  std::vector<double> fippi_records();

 private:
  const Geometry &G_;
  std::vector<DevSurface> ds_;
  std::vector<NearBlock> blocks_;
  std::vector<FippiJob> jobs_;
  std::vector<Key15> job_keys_;
  std::vector<int> mates_;
  FippiJob *d_jobs_ = nullptr;
  double *d_records_ = nullptr;
  bool fippi_ready_ = false;
  double *d_HP_ = nullptr, *d_HN_ = nullptr, *d_HS_ = nullptr, *d_scratch_ = nullptr, *d_stage_ = nullptr;
  size_t cap_HP_ = 0, cap_HN_ = 0, cap_HS_ = 0, cap_scratch_ = 0, cap_stage_ = 0;
  RegionK *d_regs_ = nullptr;
  SWJob *d_swjobs_ = nullptr;
  size_t cap_swjobs_ = 0;
  AsmStats stats_;
  std::string err_;
  std::vector<hipEvent_t> evpool_;   // timing events of the band loop
  int region_only_ = -1;
  int far_bits_ = 64;
  unsigned long long serial_ = 0, alloc_gen_ = 0;
  RegionK *d_regtab_ = nullptr;      // [frequency][block][2] of a prepared sweep
  size_t cap_regtab_ = 0;
  std::vector<int> nreg_tab_;        // [frequency][block]
  std::vector<VGSurf> vgs_;          // vertex-gradient views (vg_tables)
  std::vector<void *> vg_alloc_;
  std::vector<int *> vg_ev_;
  long vg_slots_ = 0;
  void build_near();
  void build_bands(NearBlock &B);
  // far records and matrix entries of block B, band by band (the near,
  // self and short-wavelength records must be in place); adds the far and
  // stamp times to *t_far / *t_stamp when they are given
  int far_and_stamp(NearBlock &B, const DevSurface &Da, const DevSurface &Db, const Surface &Sa,
                    const Surface &Sb, int nreg, const RegionK *dreg, int needH1, MatOut M, long rowOff,
                    long colOff, hipStream_t stream, double *t_far, double *t_stamp);
  int regions_for(int sa, int sb, const std::vector<cdouble> &eps, const std::vector<cdouble> &mu, cdouble omega,
                  RegionK rk[2]);
  int ensure(double **p, size_t *cap, size_t need);
  // short-wavelength pairs of block bi at this frequency, launched after
  // the near (and self) kernels; -5 for a pair with no defined value
  int run_sw(int sa, int sb, int bi, const RegionK *rk, int nreg, const RegionK *dreg, int needH1,
             hipStream_t stream);
};

// relative centroid distance and common-vertex count of two panels, the
// common vertices first in ia / ib (AssessPanelPair.cc:52-172)
int assess_pair(const Surface &Sa, int pa, const Surface &Sb, int pb, double *rRel, int ia[3], int ib[3]);

}  // namespace scuffx
