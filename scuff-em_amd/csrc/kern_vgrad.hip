// kern_vgrad.hip -- vertex-position derivatives of the BEM matrix, the
// plane-wave RHS and the absorbed power, contracted in the kernels.
//
// Matrix.  An entry is a sum over panel pairs (P, Q) and common regions of
//   L_a L_b s_a s_b [pf1|pf3] H0(P,qa,Q,qb)  and  L_a L_b s_a s_b pf2/(ik) H1
// (EdgeEdgeInteractions.cc:154-188, SurfaceSurfaceInteractions.cc:162-327),
// with H0 = sum w [(x-Qa).(y-Qb) + 4/(ik)^2] Phi(|x-y|),
// H1 = sum w ((x-Qa) x (y-Qb)).(x-y) Psi, Psi = Phi (ik - 1/r)/r, the
// sums over the points x = sum_k lambda_k V_k of P and y of Q.  In
// barycentric form the panel areas cancel, so an entry depends on the
// vertices through the points, the source vertices Qa, Qb and the edge
// lengths.  With the weights S = C + C^T gathered per panel pair,
//   T(P, Q) = sum_{qa,qb} W0 H0 + W1 H1
//           = sum_ij w_ij [P0(x_i, y_j) Phi_ij + P1(x_i, y_j) Psi_ij],
// P0, P1 polynomials of the points whose coefficients are weight sums.
// Its derivative with respect to vertex k of P is
//   sum_ij w_ij lambda_k(i) grad_x F_ij  +  (terms of Qa = V_k),
// evaluated point by point; no derivative block is stored.  A far pair
// uses the degree-4 rule of the assembly (its derivative is the exact
// derivative of the assembled value); a near pair without a common vertex
// uses the 78-point rule on each quarter of both panels; touching and self
// pairs come from central
// differences of their Taylor-Duffy integrals (kernels of kern_sw.hip and
// kern_self.hip evaluated at displaced vertices).  The edge lengths enter
// through the assembled matrix: sum_ij C_ij M_ij (d ln L_i + d ln L_j).
//
// RHS.  b_e = -/+ L_e sum_m w_m (X_m - Q).F(X_m) per panel, F the incident
// E / Z0 or H field (AssembleRHSVector.cc:190-246); its derivative follows
// the points, Q and L_e.
//
// Absorbed power.  PAbs = sum W_ab Re(conj(K_a) N_b - conj(N_a) K_b) with
// W_ab = -Z0/4 sum over shared panels of the crossed overlap
// L_a L_b tf / (2A) (OPFT.cc:82-146); the overlap is differentiated in
// dual numbers, and gamma collects dPAbs / dKN.
#include <hip/hip_runtime.h>

#include "vgrad.h"
#include "tcr20.inc"

namespace scuffx {

#define VG_INV4PI 0.079577471545947667884
#define VG_ZVAC 376.73031346177

__constant__ double c_vg_tcr4[18] = SCUFFX_TCR4_INIT;
__constant__ double c_vg_tcr20[234];
// the 78-point rule on each quarter of the triangle (midpoint subdivision):
// 312 points, barycentric (u, v) of the whole triangle, weights / 4
__constant__ double c_vg_tcr20q[936];

__device__ inline cplx ld2(const double2 *p, long i)
{
  const double2 z = p[i];
  return mk(z.x, z.y);
}

// S_ij = C_ij + C_ji
__device__ inline cplx wS(const VGW &w, long i, long j)
{
  if (w.C) return ld2(w.C, i + j * w.ldc) + ld2(w.C, j + i * w.ldc);
  return ld2(w.u, i) * ld2(w.v, j) + ld2(w.v, i) * ld2(w.u, j);
}

__device__ inline void atomic_c(double2 *g, long slot, cplx v)
{
  atomicAdd(&g[slot].x, v.r);
  atomicAdd(&g[slot].y, v.i);
}

// weights of the 9 (qa, qb) combinations of panel pair (p on A, q on B) in
// one region; scale 1 for distinct panels, 1/2 for a panel with itself
__device__ void pair_weights(const VGSurf &A, int p, const VGSurf &B, int q, const VGW &w, const RegionK &rk,
                             double scale, cplx W0[9], cplx W1[9])
{
  const cplx ik = IMUL(rk.k);
  const bool kzero = (rk.k.r == 0.0 && rk.k.i == 0.0);
  const cplx oik = kzero ? mk(0, 0) : 1.0 / ik;
  for (int qa = 0; qa < 3; qa++)
    for (int qb = 0; qb < 3; qb++) {
      const int ea = A.PBF[3 * p + qa], eb = B.PBF[3 * q + qb];
      W0[3 * qa + qb] = mk(0, 0);
      W1[3 * qa + qb] = mk(0, 0);
      if (ea < 0 || eb < 0 || kzero) continue;
      const double ss = scale * A.PSG[3 * p + qa] * B.PSG[3 * q + qb] * A.EL[ea] * B.EL[eb];
      const long ra = A.off, cb = B.off;
      if (A.pec && B.pec) {
        W0[3 * qa + qb] = ss * (rk.pf1 * wS(w, ra + ea, cb + eb));
      } else if (A.pec) {
        W0[3 * qa + qb] = ss * (rk.pf1 * wS(w, ra + ea, cb + 2 * eb));
        W1[3 * qa + qb] = ss * (rk.pf2 * oik * wS(w, ra + ea, cb + 2 * eb + 1));
      } else if (B.pec) {
        W0[3 * qa + qb] = ss * (rk.pf1 * wS(w, ra + 2 * ea, cb + eb));
        W1[3 * qa + qb] = ss * (rk.pf2 * oik * wS(w, ra + 2 * ea + 1, cb + eb));
      } else {
        W0[3 * qa + qb] = ss * (rk.pf1 * wS(w, ra + 2 * ea, cb + 2 * eb) + rk.pf3 * wS(w, ra + 2 * ea + 1, cb + 2 * eb + 1));
        W1[3 * qa + qb] = ss * (rk.pf2 * oik * (wS(w, ra + 2 * ea, cb + 2 * eb + 1) + wS(w, ra + 2 * ea + 1, cb + 2 * eb)));
      }
    }
}

__device__ inline void rcross(const double *a, const cplx *b, cplx *c)   // real x complex
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline void ccross(const cplx *a, const cplx *b, cplx *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// One panel pair in one region: adds to G[k][c] the derivative of
// sum_{qa,qb} W0 H0 + W1 H1 with respect to coordinate c of vertex slot k
// of the row panel.  Pp, Pq: vertices relative to the panels' own
// origins, D = origin(P) - origin(Q).  Point pairs t0, t0 + tstep, ...
// of the nq x nq rule.
__device__ void vg_pair(const double *Pp, const double *Pq, const double *D, const cplx W0[9], const cplx W1[9],
                        cplx k, int needH1, const double *rule, int nq, int t0, int tstep, cplx G[3][3])
{
#pragma clang fp contract(fast)
  const cplx ik = IMUL(k);
  const bool kzero = (k.r == 0.0 && k.i == 0.0);
  if (kzero) return;
  const cplx c4 = 4.0 / (ik * ik);
  // weight sums: P0 = w0s x.y - x.A0 - y.B0 + c0, V = w1s x*y - x*A1 - B1*y + D1
  cplx w0s = mk(0, 0), c0 = mk(0, 0), w1s = mk(0, 0);
  cplx A0[3], B0[3], A1[3], B1[3], D1[3];
  for (int c = 0; c < 3; c++) { A0[c] = mk(0, 0); B0[c] = mk(0, 0); A1[c] = mk(0, 0); B1[c] = mk(0, 0); D1[c] = mk(0, 0); }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const cplx u0 = W0[3 * a + b], u1 = W1[3 * a + b];
      const double *qa = Pp + 3 * a, *qb = Pq + 3 * b;
      w0s += u0;
      w1s += u1;
      c0 += dot3(qa, qb) * u0;
      double qq[3];
      cross3(qa, qb, qq);
      for (int c = 0; c < 3; c++) {
        A0[c] += qb[c] * u0; B0[c] += qa[c] * u0;
        A1[c] += qb[c] * u1; B1[c] += qa[c] * u1; D1[c] += qq[c] * u1;
      }
    }
  c0 += c4 * w0s;
  cplx S1 = mk(0, 0), Sy[3], Tyr[3], Tr[3], Gl[3][3];
  for (int c = 0; c < 3; c++) {
    Sy[c] = mk(0, 0); Tyr[c] = mk(0, 0); Tr[c] = mk(0, 0);
    for (int kk = 0; kk < 3; kk++) Gl[kk][c] = mk(0, 0);
  }
  const int ntot = nq * nq;
  for (int t = t0; t < ntot; t += tstep) {
    const int i = t / nq, j = t - i * nq;
    const double ui = rule[3 * i], vi = rule[3 * i + 1], uj = rule[3 * j], vj = rule[3 * j + 1];
    const double w = rule[3 * i + 2] * rule[3 * j + 2];
    const double lam[3] = {1.0 - ui - vi, ui, vi};
    double x[3], y[3], R[3];
    for (int c = 0; c < 3; c++) {
      x[c] = lam[0] * Pp[c] + ui * Pp[3 + c] + vi * Pp[6 + c];
      y[c] = (1.0 - uj - vj) * Pq[c] + uj * Pq[3 + c] + vj * Pq[6 + c];
      R[c] = D[c] + x[c] - y[c];
    }
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]), ir = 1.0 / r;
    double e = exp(-k.i * r), sn, cs;
    sincos(k.r * r, &sn, &cs);
    const cplx Phi = mk(e * cs, e * sn) * (VG_INV4PI * ir);
    const cplx g = (ik - mk(ir, 0)) * ir;                  // Psi = Phi g
    const cplx gp = mk(2.0 * ir * ir * ir, 0) - ik * (ir * ir);   // dg/dr
    const cplx Psi = Phi * g;
    const cplx Psip = Phi * (g * g * r + gp);              // dPsi/dr
    const double xy = dot3(x, y);
    cplx P0 = w0s * xy + c0;
    for (int c = 0; c < 3; c++) P0 = P0 - x[c] * A0[c] - y[c] * B0[c];
    cplx gr[3];
    const cplx P0Psi = P0 * Psi;
    for (int c = 0; c < 3; c++) gr[c] = (w0s * y[c] - A0[c]) * Phi + R[c] * P0Psi;
    if (needH1) {
      double xxy[3];
      cross3(x, y, xxy);
      cplx xA1[3], yB1[3], V[3], U[3], UR[3];
      rcross(x, A1, xA1);
      rcross(y, B1, yB1);   // y x B1 = -(B1 x y)
      for (int c = 0; c < 3; c++) {
        V[c] = w1s * xxy[c] - xA1[c] + yB1[c] + D1[c];
        U[c] = w1s * y[c] - A1[c];
      }
      // U x R = -(R x U)
      rcross(R, U, UR);
      const cplx P1 = V[0] * R[0] + V[1] * R[1] + V[2] * R[2];
      const cplx fac = P1 * Psip * ir;
      for (int c = 0; c < 3; c++) gr[c] += (V[c] - UR[c]) * Psi + R[c] * fac;
      double yR[3];
      cross3(y, R, yR);
      const cplx wPsi = w * Psi;
      for (int c = 0; c < 3; c++) { Tyr[c] += yR[c] * wPsi; Tr[c] += R[c] * wPsi; }
    }
    const cplx wPhi = w * Phi;
    S1 += wPhi;
    for (int c = 0; c < 3; c++) Sy[c] += y[c] * wPhi;
    for (int kk = 0; kk < 3; kk++) {
      const double wl = w * lam[kk];
      for (int c = 0; c < 3; c++) Gl[kk][c] += wl * gr[c];
    }
  }
  // source-vertex terms: d/dQa of the combinations with qa = k
  for (int kk = 0; kk < 3; kk++) {
    cplx a0 = mk(0, 0), a1 = mk(0, 0), e0[3], e1[3], ex[3];
    for (int c = 0; c < 3; c++) { e0[c] = mk(0, 0); e1[c] = mk(0, 0); }
    for (int b = 0; b < 3; b++) {
      const cplx u0 = W0[3 * kk + b], u1 = W1[3 * kk + b];
      a0 += u0; a1 += u1;
      for (int c = 0; c < 3; c++) { e0[c] += Pq[3 * b + c] * u0; e1[c] += Pq[3 * b + c] * u1; }
    }
    ccross(e1, Tr, ex);
    for (int c = 0; c < 3; c++)
      G[kk][c] += Gl[kk][c] - a0 * Sy[c] + e0[c] * S1 - a1 * Tyr[c] + ex[c];
  }
}

// panel p's vertices relative to its centroid
__device__ inline void panel_local(const VGSurf &A, int p, double *O, double *P)
{
  const double *V0 = A.V + 3 * A.PVI[3 * p], *V1 = A.V + 3 * A.PVI[3 * p + 1], *V2 = A.V + 3 * A.PVI[3 * p + 2];
  for (int c = 0; c < 3; c++) {
    O[c] = (V0[c] + V1[c] + V2[c]) / 3.0;
    P[c] = V0[c] - O[c]; P[3 + c] = V1[c] - O[c]; P[6 + c] = V2[c] - O[c];
  }
}

__device__ inline int vg_near_find(const int *rowPtr, const int *col, int pa, int pb)
{
  int lo = rowPtr[pa], hi = rowPtr[pa + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1, c = col[mid];
    if (c == pb) return mid;
    if (c < pb) lo = mid + 1; else hi = mid;
  }
  return -1;
}

// block of 256 per row panel; the 18 partial sums of the block reduced
// through shuffles and LDS, then added to the panel's three vertices
__global__ void __launch_bounds__(256)
k_vg_far(VGSurf A, VGSurf B, int same, const int *nearRowPtr, const int *nearCol, int nreg, const RegionK *regs,
         int needH1, VGW w, double2 *grad)
{
  __shared__ double red[4][18];
  const int p = blockIdx.x;
  double Op[3], Pp[9];
  panel_local(A, p, Op, Pp);
  cplx G[3][3];
  for (int kk = 0; kk < 3; kk++) for (int c = 0; c < 3; c++) G[kk][c] = mk(0, 0);
  for (int q = threadIdx.x; q < B.nP; q += blockDim.x) {
    if (same && q == p) continue;
    if (vg_near_find(nearRowPtr, nearCol, p, q) >= 0) continue;
    double Oq[3], Pq[9], D[3];
    panel_local(B, q, Oq, Pq);
    for (int c = 0; c < 3; c++) D[c] = Op[c] - Oq[c];
    for (int r = 0; r < nreg; r++) {
      cplx W0[9], W1[9];
      pair_weights(A, p, B, q, w, regs[r], 1.0, W0, W1);
      vg_pair(Pp, Pq, D, W0, W1, regs[r].k, needH1, c_vg_tcr4, 6, 0, 1, G);
    }
  }
  double v[18];
  for (int kk = 0; kk < 3; kk++)
    for (int c = 0; c < 3; c++) { v[6 * kk + 2 * c] = G[kk][c].r; v[6 * kk + 2 * c + 1] = G[kk][c].i; }
  for (int off = 32; off > 0; off >>= 1)
    for (int m = 0; m < 18; m++) v[m] += __shfl_xor(v[m], off);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int m = 0; m < 18; m++) red[wave][m] = v[m];
  __syncthreads();
  if (threadIdx.x < 18) {
    const int m = threadIdx.x;
    const double s = red[0][m] + red[1][m] + red[2][m] + red[3][m];
    const int kk = m / 6, c = (m % 6) / 2, im = m & 1;
    const long slot = A.gofs + 3L * A.PVI[3 * p + kk] + c;
    atomicAdd(im ? &grad[slot].y : &grad[slot].x, s);
  }
}

// one wave per listed near pair, lanes over the 312 x 312 point pairs of
// the subdivided 78-point rule (the derivative's integrand falls off as
// 1/r^3 at these distances: one 78-point rule per panel leaves ~1e-5 of
// it for the closest pairs)
__global__ void __launch_bounds__(64)
k_vg_nearsep(VGSurf A, VGSurf B, const int *pairs, int npairs, int nreg, const RegionK *regs, int needH1, VGW w,
             double2 *grad)
{
  const int t = blockIdx.x;
  if (t >= npairs) return;
  const int p = pairs[2 * t], q = pairs[2 * t + 1];
  double Op[3], Pp[9], Oq[3], Pq[9], D[3];
  panel_local(A, p, Op, Pp);
  panel_local(B, q, Oq, Pq);
  for (int c = 0; c < 3; c++) D[c] = Op[c] - Oq[c];
  cplx G[3][3];
  for (int kk = 0; kk < 3; kk++) for (int c = 0; c < 3; c++) G[kk][c] = mk(0, 0);
  for (int r = 0; r < nreg; r++) {
    cplx W0[9], W1[9];
    pair_weights(A, p, B, q, w, regs[r], 1.0, W0, W1);
    vg_pair(Pp, Pq, D, W0, W1, regs[r].k, needH1, c_vg_tcr20q, 312, threadIdx.x, 64, G);
  }
  double v[18];
  for (int kk = 0; kk < 3; kk++)
    for (int c = 0; c < 3; c++) { v[6 * kk + 2 * c] = G[kk][c].r; v[6 * kk + 2 * c + 1] = G[kk][c].i; }
  for (int off = 32; off > 0; off >>= 1)
    for (int m = 0; m < 18; m++) v[m] += __shfl_xor(v[m], off);
  if (threadIdx.x == 0)
    for (int kk = 0; kk < 3; kk++) {
      const long slot = A.gofs + 3L * A.PVI[3 * p + kk];
      for (int c = 0; c < 3; c++) {
        atomicAdd(&grad[slot + c].x, v[6 * kk + 2 * c]);
        atomicAdd(&grad[slot + c].y, v[6 * kk + 2 * c + 1]);
      }
    }
}

__global__ void k_vg_fd(const VGSurf *surfs, const VGFDItem *items, int nitems, const RegionK *regs,
                        const double *HN, const double *HS, VGW w, double2 *grad)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nitems) return;
  const VGFDItem it = items[t];
  const VGSurf &A = surfs[it.sa], &B = surfs[it.sb];
  const double *Hp = (it.self ? HS : HN) + it.hofs;
  const long half = (long)it.nreg * 9 * it.stride;
  const double *Hm = Hp + half;
  const double i2h = 0.5 / it.h;
  cplx acc = mk(0, 0);
  for (int r = 0; r < it.nreg; r++) {
    cplx W0[9], W1[9];
    pair_weights(A, it.pa, B, it.pb, w, regs[2 * it.blk + r], it.self ? 0.5 : 1.0, W0, W1);
    for (int q = 0; q < 9; q++) {
      const double *a = Hp + (r * 9 + q) * it.stride, *b = Hm + (r * 9 + q) * it.stride;
      acc += W0[q] * mk((a[0] - b[0]) * i2h, (a[1] - b[1]) * i2h);
      if (it.stride == 4) acc += W1[q] * mk((a[2] - b[2]) * i2h, (a[3] - b[3]) * i2h);
    }
  }
  atomic_c(grad, it.slot, acc);
}

// rows: rho_i += sum_j C_ij M_ij over a chunk of columns
__global__ void __launch_bounds__(256)
k_vg_rho_rows(const double2 *M, long ldm, int n, VGW w, int chunk, double2 *rho)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  cplx acc = mk(0, 0);
  if (w.C) {
    for (int j = j0; j < j1; j++) acc += ld2(w.C, i + (long)j * w.ldc) * ld2(M, i + (long)j * ldm);
  } else {
    for (int j = j0; j < j1; j++) acc += ld2(w.v, j) * ld2(M, i + (long)j * ldm);
    acc = acc * ld2(w.u, i);
  }
  atomic_c(rho, i, acc);
}

// columns: rho_i += sum_j C_ji M_ji, one block per column i
__global__ void __launch_bounds__(256)
k_vg_rho_cols(const double2 *M, long ldm, int n, VGW w, double2 *rho)
{
  __shared__ double sr[4], si[4];
  const int i = blockIdx.x;
  cplx acc = mk(0, 0);
  if (w.C) {
    for (int j = threadIdx.x; j < n; j += 256) acc += ld2(w.C, j + (long)i * w.ldc) * ld2(M, j + (long)i * ldm);
  } else {
    for (int j = threadIdx.x; j < n; j += 256) acc += ld2(w.u, j) * ld2(M, j + (long)i * ldm);
  }
  for (int off = 32; off > 0; off >>= 1) { acc.r += __shfl_xor(acc.r, off); acc.i += __shfl_xor(acc.i, off); }
  if ((threadIdx.x & 63) == 0) { sr[threadIdx.x >> 6] = acc.r; si[threadIdx.x >> 6] = acc.i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    cplx s = mk(sr[0] + sr[1] + sr[2] + sr[3], si[0] + si[1] + si[2] + si[3]);
    if (!w.C) s = s * ld2(w.v, i);
    atomic_c(rho, i, s);
  }
}

__global__ void k_vg_len(VGSurf A, const int *EV, int nE, const double2 *rho, double2 *grad)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  cplx rr = A.pec ? ld2(rho, A.off + e) : ld2(rho, A.off + 2 * e) + ld2(rho, A.off + 2 * e + 1);
  const int v1 = EV[4 * e], v2 = EV[4 * e + 1];
  const double *X1 = A.V + 3 * v1, *X2 = A.V + 3 * v2;
  double d[3];
  sub3(X2, X1, d);
  const double il2 = 1.0 / dot3(d, d);
  for (int c = 0; c < 3; c++) {
    const double g = d[c] * il2;   // d ln L / d X2_c
    atomic_c(grad, A.gofs + 3L * v2 + c, g * rr);
    atomic_c(grad, A.gofs + 3L * v1 + c, -g * rr);
  }
}

// one panel of one edge: I[0..1] = sum w (X - Q).E, I[2..3] = (X - Q).H
// (without the edge length) and its derivatives dI[slot][c] with respect to
// the panel's vertex slots (V1, V2, V3 in the call order; Q = slot qs)
__device__ void vg_rhs_panel(const double *V1, const double *V2, const double *V3, int qs, const cplx *E0,
                             const double *nh, cplx K, cplx Z, cplx I[2], cplx dI[2][3][3])
{
  // b carries pre * J = (L / 2 Area) * 2 Area = L: the panel's area does
  // not enter, and the sums run over the reference triangle's weights
  const double *Vs[3] = {V1, V2, V3};
  const double *Q = Vs[qs];
  double A[3], B[3];
  sub3(V2, V1, A); sub3(V3, V1, B);
  I[0] = mk(0, 0); I[1] = mk(0, 0);
  for (int f = 0; f < 2; f++) for (int s = 0; s < 3; s++) for (int c = 0; c < 3; c++) dI[f][s][c] = mk(0, 0);
  const cplx iK = IMUL(K);
  for (int m = 0; m < 78; m++) {
    const double u = c_vg_tcr20[3 * m], v = c_vg_tcr20[3 * m + 1], w = c_vg_tcr20[3 * m + 2];
    const double lam[3] = {1.0 - u - v, u, v};
    double X[3], f[3];
    for (int c = 0; c < 3; c++) { X[c] = V1[c] + u * A[c] + v * B[c]; f[c] = X[c] - Q[c]; }
    const cplx ef = cexp_(iK * (nh[0] * X[0] + nh[1] * X[1] + nh[2] * X[2]));
    cplx F[2][3];
    for (int c = 0; c < 3; c++) F[0][c] = E0[c] * ef;
    F[1][0] = (nh[1] * F[0][2] - nh[2] * F[0][1]) / Z;
    F[1][1] = (nh[2] * F[0][0] - nh[0] * F[0][2]) / Z;
    F[1][2] = (nh[0] * F[0][1] - nh[1] * F[0][0]) / Z;
    for (int fi = 0; fi < 2; fi++) {
      const cplx fF = f[0] * F[fi][0] + f[1] * F[fi][1] + f[2] * F[fi][2];
      I[fi] += w * fF;
      const cplx t = w * (fF * iK);
      for (int s = 0; s < 3; s++) {
        const double dl = lam[s] - (s == qs ? 1.0 : 0.0);
        for (int c = 0; c < 3; c++) dI[fi][s][c] += (w * dl) * F[fi][c] + (lam[s] * nh[c]) * t;
      }
    }
  }
}

__global__ void k_vg_rhs(VGSurf A, const int *EV, int nE, double sgn, cplx e0x, cplx e0y, cplx e0z, double nx,
                         double ny, double nz, cplx K, cplx Z, const double2 *mu, double2 *grad)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  const cplx E0[3] = {e0x, e0y, e0z};
  const double nh[3] = {nx, ny, nz};
  const int iV1 = EV[4 * e], iV2 = EV[4 * e + 1], iQP = EV[4 * e + 2], iQM = EV[4 * e + 3];
  const double *X1 = A.V + 3 * iV1, *X2 = A.V + 3 * iV2;
  double d[3];
  sub3(X2, X1, d);
  const double L = sqrt(dot3(d, d));
  // b_E = sgn L (I_E(P) - I_E(M)) / Z0, b_H = sgn L (I_H(P) - I_H(M))
  const cplx wE = A.pec ? ld2(mu, A.off + e) * (sgn / VG_ZVAC) : ld2(mu, A.off + 2 * e) * (sgn / VG_ZVAC);
  const cplx wH = A.pec ? mk(0, 0) : ld2(mu, A.off + 2 * e + 1) * sgn;
  cplx I[2], dI[2][3][3];
  cplx Itot = mk(0, 0);
  for (int side = 0; side < 2; side++) {
    int vs[3], qs;
    double ps;
    if (side == 0) { vs[0] = iQP; vs[1] = iV1; vs[2] = iV2; qs = 0; ps = 1.0; }
    else {
      if (iQM < 0) break;
      vs[0] = iV1; vs[1] = iV2; vs[2] = iQM; qs = 2; ps = -1.0;
    }
    vg_rhs_panel(A.V + 3 * vs[0], A.V + 3 * vs[1], A.V + 3 * vs[2], qs, E0, nh, K, Z, I, dI);
    Itot += ps * (wE * I[0] + wH * I[1]);
    for (int s = 0; s < 3; s++)
      for (int c = 0; c < 3; c++)
        atomic_c(grad, A.gofs + 3L * vs[s] + c, (ps * L) * (wE * dI[0][s][c] + wH * dI[1][s][c]));
  }
  // edge length
  for (int c = 0; c < 3; c++) {
    const double g = d[c] / L;
    atomic_c(grad, A.gofs + 3L * iV2 + c, g * Itot);
    atomic_c(grad, A.gofs + 3L * iV1 + c, -g * Itot);
  }
}

// forward-mode dual numbers for the overlap derivative
struct dn {
  double v, d;
};
__device__ inline dn dmk(double v, double d) { dn r; r.v = v; r.d = d; return r; }
__device__ inline dn operator+(dn a, dn b) { return dmk(a.v + b.v, a.d + b.d); }
__device__ inline dn operator-(dn a, dn b) { return dmk(a.v - b.v, a.d - b.d); }
__device__ inline dn operator*(dn a, dn b) { return dmk(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline dn operator*(double s, dn a) { return dmk(s * a.v, s * a.d); }
__device__ inline dn operator/(dn a, dn b) { return dmk(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
__device__ inline dn dsqrt(dn a) { const double s = sqrt(a.v); return dmk(s, 0.5 * a.d / s); }

// crossed overlap of vertex slots qa, qb on a panel (X: its vertices in
// mesh order), without the sign: L_a L_b tf / (2 Area), tf =
// (2 L1 + L2).(Z x (Qa - Qb)) / 6, Z the unit normal times fl
__device__ dn overlap_d(const dn X[3][3], int qa, int qb, double fl)
{
  dn Aa[3], Bb[3], N[3];
  for (int c = 0; c < 3; c++) { Aa[c] = X[1][c] - X[0][c]; Bb[c] = X[2][c] - X[0][c]; }
  N[0] = Aa[1] * Bb[2] - Aa[2] * Bb[1];
  N[1] = Aa[2] * Bb[0] - Aa[0] * Bb[2];
  N[2] = Aa[0] * Bb[1] - Aa[1] * Bb[0];
  const dn nn = dsqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);   // 2 Area
  dn Z[3];
  for (int c = 0; c < 3; c++) Z[c] = fl * (N[c] / nn);
  const dn *Qa = X[qa], *Q1 = X[(qa + 1) % 3], *Q2 = X[(qa + 2) % 3], *Qb = X[qb];
  dn L1[3], L2[3], DQ[3], ZxDQ[3];
  for (int c = 0; c < 3; c++) { L1[c] = Q1[c] - Qa[c]; L2[c] = Q2[c] - Q1[c]; DQ[c] = Qa[c] - Qb[c]; }
  ZxDQ[0] = Z[1] * DQ[2] - Z[2] * DQ[1];
  ZxDQ[1] = Z[2] * DQ[0] - Z[0] * DQ[2];
  ZxDQ[2] = Z[0] * DQ[1] - Z[1] * DQ[0];
  dn tf = dmk(0, 0);
  for (int c = 0; c < 3; c++) tf = tf + (2.0 * L1[c] + L2[c]) * ZxDQ[c];
  tf = (1.0 / 6.0) * tf;
  // edge of slot q spans the other two vertices
  auto elen = [&](int q) {
    const dn *P1 = X[(q + 1) % 3], *P2 = X[(q + 2) % 3];
    dn s = dmk(0, 0);
    for (int c = 0; c < 3; c++) { const dn t = P2[c] - P1[c]; s = s + t * t; }
    return dsqrt(s);
  };
  return elen(qa) * elen(qb) * tf / nn;
}

// thread per edge a, as the power kernel (OPFT crossed overlaps)
__global__ void k_vg_pabs(VGSurf A, const double *PZ, const int *PEI, const int *EI, int nE, const double2 *KN,
                          double *gradR, double2 *gamma)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nE) return;
  const cplx Ka = ld2(KN, A.off + 2 * a), Na = ld2(KN, A.off + 2 * a + 1);
  cplx gK = mk(0, 0), gN = mk(0, 0);
  const int pa2[2] = {EI[4 * a], EI[4 * a + 1]}, qa2[2] = {EI[4 * a + 2], EI[4 * a + 3]};
  for (int sa = 0; sa < 2; sa++) {
    const int p = pa2[sa];
    if (p < 0) continue;
    const int vi[3] = {A.PVI[3 * p], A.PVI[3 * p + 1], A.PVI[3 * p + 2]};
    double nrm[3], Av[3], Bv[3];
    sub3(A.V + 3 * vi[1], A.V + 3 * vi[0], Av);
    sub3(A.V + 3 * vi[2], A.V + 3 * vi[0], Bv);
    cross3(Av, Bv, nrm);
    const double fl = dot3(nrm, PZ + 3 * p) >= 0.0 ? 1.0 : -1.0;
    for (int le = 0; le < 3; le++) {
      const int b = PEI[3 * p + le];
      if (b < 0) continue;
      const int sb = (EI[4 * b] == p) ? 0 : 1;
      const int qb = sb == 0 ? EI[4 * b + 2] : EI[4 * b + 3], qa = qa2[sa];
      const double sign = (sa == sb) ? 1.0 : -1.0;
      const cplx Kb = ld2(KN, A.off + 2 * b), Nb = ld2(KN, A.off + 2 * b + 1);
      // re_ab = -Z0 Re(conj(Ka) Nb - conj(Na) Kb), PAbs += re_ab ox_ab / 4
      const double re = -VG_ZVAC * ((Ka.r * Nb.r + Ka.i * Nb.i) - (Na.r * Kb.r + Na.i * Kb.i));
      dn X[3][3];
      for (int s = 0; s < 3; s++) for (int c = 0; c < 3; c++) X[s][c] = dmk(A.V[3 * vi[s] + c], 0.0);
      const double oab = sign * overlap_d(X, qa, qb, fl).v, oba = sign * overlap_d(X, qb, qa, fl).v;
      // gamma: coefficients of dK_a and dN_a, W_xy = -Z0/4 ox_xy
      const double wab = -0.25 * VG_ZVAC * oab, wba = -0.25 * VG_ZVAC * oba;
      gK += (wab - wba) * mk(Nb.r, -Nb.i);
      gN += (wba - wab) * mk(Kb.r, -Kb.i);
      for (int s = 0; s < 3; s++)
        for (int c = 0; c < 3; c++) {
          X[s][c].d = 1.0;
          const double dox = sign * overlap_d(X, qa, qb, fl).d;
          X[s][c].d = 0.0;
          atomicAdd(&gradR[A.gofs + 3L * vi[s] + c], 0.25 * re * dox);
        }
    }
  }
  gamma[A.off + 2 * a] = make_double2(gK.r, gK.i);
  gamma[A.off + 2 * a + 1] = make_double2(gN.r, gN.i);
}

static bool g_vg_ready[64] = {false};

static hipError_t vg_tables()
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (g_vg_ready[dev & 63]) return hipSuccess;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_vg_tcr20), scuffx_tcr20, sizeof(scuffx_tcr20));
  if (e != hipSuccess) return e;
  // quarters: corners (a, b, c) in (u, v) of the whole triangle
  static const double Q[4][3][2] = {{{0, 0}, {0.5, 0}, {0, 0.5}},
                                    {{0.5, 0}, {1, 0}, {0.5, 0.5}},
                                    {{0, 0.5}, {0.5, 0.5}, {0, 1}},
                                    {{0.5, 0.5}, {0, 0.5}, {0.5, 0}}};
  static double hq[936];
  for (int t = 0; t < 4; t++)
    for (int i = 0; i < 78; i++) {
      const double u = scuffx_tcr20[3 * i], v = scuffx_tcr20[3 * i + 1], w = scuffx_tcr20[3 * i + 2];
      double *o = hq + 3 * (78 * t + i);
      for (int c = 0; c < 2; c++) o[c] = Q[t][0][c] + u * (Q[t][1][c] - Q[t][0][c]) + v * (Q[t][2][c] - Q[t][0][c]);
      o[2] = 0.25 * w;
    }
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_vg_tcr20q), hq, sizeof(hq));
  if (e == hipSuccess) g_vg_ready[dev & 63] = true;
  return e;
}

hipError_t launch_vg_far(hipStream_t s, const VGSurf &A, const VGSurf &B, int same, const int *nearRowPtr,
                         const int *nearCol, int nreg, const RegionK *regs, int needH1, const VGW &w,
                         double2 *grad)
{
  if (A.nP <= 0 || B.nP <= 0 || nreg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vg_far, dim3(A.nP), dim3(256), 0, s, A, B, same, nearRowPtr, nearCol, nreg, regs, needH1, w,
                     grad);
  return hipGetLastError();
}

hipError_t launch_vg_nearsep(hipStream_t s, const VGSurf &A, const VGSurf &B, const int *pairs, int npairs,
                             int nreg, const RegionK *regs, int needH1, const VGW &w, double2 *grad)
{
  if (npairs <= 0 || nreg <= 0) return hipSuccess;
  hipError_t e = vg_tables();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_vg_nearsep, dim3(npairs), dim3(64), 0, s, A, B, pairs, npairs, nreg, regs, needH1, w, grad);
  return hipGetLastError();
}

hipError_t launch_vg_fd(hipStream_t s, const VGSurf *surfs, const VGFDItem *items, int nitems,
                        const RegionK *regs, const double *HN, const double *HS, const VGW &w, double2 *grad)
{
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vg_fd, dim3((nitems + 127) / 128), dim3(128), 0, s, surfs, items, nitems, regs, HN, HS, w,
                     grad);
  return hipGetLastError();
}

hipError_t launch_vg_rho(hipStream_t s, const double2 *M, long ldm, int n, const VGW &w, double2 *rho)
{
  if (n <= 0) return hipSuccess;
  const int chunk = 128;
  hipLaunchKernelGGL(k_vg_rho_rows, dim3((n + 255) / 256, (n + chunk - 1) / chunk), dim3(256), 0, s, M, ldm, n, w,
                     chunk, rho);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_vg_rho_cols, dim3(n), dim3(256), 0, s, M, ldm, n, w, rho);
  return hipGetLastError();
}

hipError_t launch_vg_len(hipStream_t s, const VGSurf &A, const int *EV, int nE, const double2 *rho, double2 *grad)
{
  if (nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vg_len, dim3((nE + 127) / 128), dim3(128), 0, s, A, EV, nE, rho, grad);
  return hipGetLastError();
}

hipError_t launch_vg_rhs(hipStream_t s, const VGSurf &A, const int *EV, int nE, double sgn, const cdouble *E0,
                         const double *nhat, cdouble K, cdouble Z, const double2 *mu, double2 *grad)
{
  if (nE <= 0) return hipSuccess;
  hipError_t e = vg_tables();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_vg_rhs, dim3((nE + 63) / 64), dim3(64), 0, s, A, EV, nE, sgn,
                     mk(E0[0].real(), E0[0].imag()), mk(E0[1].real(), E0[1].imag()), mk(E0[2].real(), E0[2].imag()),
                     nhat[0], nhat[1], nhat[2], mk(K.real(), K.imag()), mk(Z.real(), Z.imag()), mu, grad);
  return hipGetLastError();
}

hipError_t launch_vg_pabs(hipStream_t s, const VGSurf &A, const double *PZ, const int *PEI, const int *EI, int nE,
                          const double2 *KN, double *gradR, double2 *gamma)
{
  if (nE <= 0 || A.pec) return hipSuccess;
  hipLaunchKernelGGL(k_vg_pabs, dim3((nE + 63) / 64), dim3(64), 0, s, A, PZ, PEI, EI, nE, KN, gradR, gamma);
  return hipGetLastError();
}

}  // namespace scuffx
