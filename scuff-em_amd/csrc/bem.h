// bem.h -- device data layout and kernel entry points of the BEM matrix
// assembly (host orchestration in assemble.cpp).
//
// Layout in HBM, per surface:
//   V      double[3*nV]                vertex coordinates
//   PVI    int[3*nP]                   panel vertex indices (mesh order)
//   PC     double[4*nP]                centroid xyz, bounding radius
//   EI     int[4*nE]                   iPPanel, iMPanel, PIndex, MIndex
//   EL     double[nE]                  edge length
// Panel-pair integrals, one record per (panel pair, region):
//   far    per column band of the block: ordered pairs (pa, pb) of every
//          row panel pa < rowLimit with the band's column panels (slots),
//          HB[slot][pa]; 9 (Qa,Qb) combinations x {H0, H1} complex per
//          region = 36 doubles (H1 omitted for PEC-PEC blocks)
//   near   ordered pairs (pa as the row panel), same record layout
//   self   panel p with itself, 9 combinations, H0 only
// The matrix M is column-major complex128, M[r + c*ldm].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "devmath.h"

namespace scuffx {

// one unique frequency-independent panel-pair integral job
struct FippiJob {
  double OVa[9];      // canonically ordered vertices of panel a
  double OVb[9];      // canonically ordered vertices of panel b
  int ncv;
  int pad;
};
#define SCUFFX_QIF_LEN 36   // xMxpRM3[3] xXxpRM3[3] RM1[9] R1[9] R2[9] x0[3]

// one ordered near pair (0 < ncv < 3 or ncv == 0 with rRel <= 4)
struct NearPair {
  double Va[9];       // vertices ordered common-first (AssessPanelPair)
  double Vb[9];
  double Qa[9];       // panel a vertices in mesh order (Q candidates)
  double Qb[9];
  double OVa0[3], OVa1[3], OVa2[3];   // canonical order (for QDFIPPI)
  double OVb0[3], OVb1[3], OVb2[3];
  int flipped;
  int ncv;
  int rec;            // FIPPI record index
  int pad;
};

// destination of the stamped matrix entries: complex128 (M[r + c ld]),
// real part only (M[r + c ld], the LHM_REAL matrix of an imaginary
// frequency) or real part in packed upper storage (M[r + c (c+1)/2] for
// r <= c, LHM_SYMMETRIC; entries below the diagonal are not stored)
enum { MAT_C128 = 0, MAT_F64 = 1, MAT_F64_PACKED = 2 };
struct MatOut {
  void *p;
  long ld;
  int kind;
  __device__ inline void put(long r, long c, cplx v) const
  {
    if (kind == MAT_C128) ((double2 *)p)[r + c * ld] = make_double2(v.r, v.i);
    else if (kind == MAT_F64) ((double *)p)[r + c * ld] = v.r;
    else if (r <= c) ((double *)p)[r + c * (c + 1) / 2] = v.r;
  }
};

struct RegionK {
  cplx k;             // wavenumber in the region
  cplx pf1, pf2, pf3; // i*mu*w*s, -i*k*s, -i*eps*w*s
};

// launch wrappers (kern_*.hip); all asynchronous on `stream`
// far records of one column band: rows pa < rowLimit x the band's slot panels
// (fp32: the point sums in single precision, records still double)
hipError_t launch_far(hipStream_t stream, const double *Va, const int *PVIa, const double *PCa, int rowLimit,
                      const double *Vb, const int *PVIb, const double *PCb, const int *slotPanels, int nslots,
                      int same, const int *nearRowPtr, const int *nearCol, int nreg, const RegionK *regs,
                      int needH1, double *HB, long rowsCap, int fp32 = 0);
hipError_t launch_near(hipStream_t stream, const NearPair *np, int nnear, const double *records,
                       int nreg, const RegionK *regs, int needH1, double *HN);
hipError_t launch_self(hipStream_t stream, const double *V, const int *PVI, int nP, int nreg,
                       const RegionK *regs, double *HS, double *scratch, int chunk);
hipError_t launch_fippi(hipStream_t stream, const FippiJob *jobs, int njobs, double *records,
                        double *heapData, int *heapSlot, double *heapKey, int cap, int nslots,
                        int *counter, int *overflow, const int *jobList);
// matrix entries of one column band: its column edges bandCols[0..ncols),
// the slots of their two panels colSlots[2 j + t], row edges < rowEdges
hipError_t launch_stamp(hipStream_t stream, int nEa, const int *EIa, const double *ELa, const int *bandCols,
                        const int *colSlots, int ncols, int rowEdges, const int *EIb, const double *ELb, int same,
                        int pecA, int pecB, int nreg, const RegionK *regs, int needH1,
                        const double *HB, long rowsCap, const double *HN, const double *HS,
                        const int *nearRowPtr, const int *nearCol, const int *nearIdx,
                        MatOut M, int rowOff, int colOff);
hipError_t launch_mirror(hipStream_t stream, MatOut M, int n);
hipError_t launch_copy_block(hipStream_t stream, MatOut M, long from, long to, long n);
// short-wavelength panel pairs (kern_sw.hip): kind 0 = near pair without a
// common vertex (all 9 combinations, 78-point rules), 1 = common edge /
// vertex near pair (Taylor-Duffy, combination q); pair = near-pair index
struct SWJob {
  int pair, region, q, kind;
};
// doubles of scratch per Taylor-Duffy job (its cached integrand values)
#define SCUFFX_SW_SLAB (2112 * 6)
hipError_t launch_sw(hipStream_t stream, const SWJob *jobs, int njobs, const NearPair *nps, const double *V,
                     const int *PVI, int nreg, const RegionK *regs, int needH1, double *HN, double *HS,
                     double *scratch, int chunk);
// rotation generators of up to three torque axes, Gamma[9 t + i + 3 j]
struct GammaSet {
  int n;
  double g[27];
};
hipError_t launch_grad(hipStream_t stream, int nEa, const int *EIa, const double *ELa, const double *Va,
                       const int *PVIa, const double *PCa, int nEb, const int *EIb, const double *ELb,
                       const double *Vb, const int *PVIb, const double *PCb, int pecA, int pecB, int nreg,
                       const RegionK *regs, double2 *GB, long ldg, long gstride, const GammaSet &GS,
                       double2 *TB);

}  // namespace scuffx
