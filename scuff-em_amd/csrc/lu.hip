// lu.hip -- blocked complex LU with partial pivoting, triangular solves
// and the refinement helpers, complex128 and complex64.
//
// Right-looking blocked algorithm (outer block nb, inner block 32):
//  * a 32-column sub-panel is factored column by column by one
//    cooperative launch: every workgroup owns a slice of rows, finds its
//    largest |re|+|im| entry, a grid barrier picks the global pivot (first
//    maximal row, as izamax), the two rows are exchanged through a small
//    global buffer, and each workgroup scales and rank-1 updates its rows;
//  * row interchanges are applied to the other columns (laswp),
//  * U12 = L11^{-1} A12 (unit-lower triangular solve),
//  * A22 -= A21 U12 on the matrix cores: v_mfma_f64_16x16x4f64 for
//    complex128 and v_mfma_f32_16x16x4f32 for complex64, four real
//    products per complex product.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "lu.h"
#include "prof.h"
#include "ws.h"

namespace cg = cooperative_groups;

namespace scuffx {


typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct CT;
// Both 16x16x4 instructions take A[i][k] from lane i + 16k and B[k][j]
// from lane j + 16k; result element r of a lane sits in column lane % 16
// and row 4*(lane/16) + r for f32 but row lane/16 + 4r for f64 (measured
// on gfx950 with a one-wave probe).
template <> struct CT<double2> {
  typedef double R;
  typedef d4 V4;
  __device__ static inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static inline V4 zero() { return V4{0, 0, 0, 0}; }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct CT<double> {
  typedef double R;
  typedef d4 V4;
  __device__ static inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static inline V4 zero() { return V4{0, 0, 0, 0}; }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct CT<float2> {
  typedef float R;
  typedef f4 V4;
  __device__ static inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static inline V4 zero() { return V4{0, 0, 0, 0}; }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T> __device__ inline T cmul(T a, T b) { T r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r; }
template <typename T> __device__ inline T csub(T a, T b) { T r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T> __device__ inline T cadd(T a, T b) { T r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T> __device__ inline T crecip(T a)
{
  // Smith's algorithm for 1/a
  T r;
  typedef decltype(a.x) R;
  if (fabs(a.x) >= fabs(a.y)) { R t = a.y / a.x, d = a.x + a.y * t; r.x = 1 / d; r.y = -t / d; }
  else { R t = a.x / a.y, d = a.x * t + a.y; r.x = t / d; r.y = -1 / d; }
  return r;
}
// real (double) counterparts, so that the panel and the solves also
// instantiate for real matrices (LHM_REAL, dgetrf/dgetrs semantics)
__device__ inline double cmul(double a, double b) { return a * b; }
__device__ inline double csub(double a, double b) { return a - b; }
__device__ inline double cadd(double a, double b) { return a + b; }
__device__ inline double crecip(double a) { return 1.0 / a; }
__device__ inline double cdiv(double a, double b) { return a / b; }
// pivot magnitude |re| + |im| (izamax / idamax)
__device__ inline double mag1(double a) { return fabs(a); }
template <typename T> __device__ inline auto mag1(T a) -> decltype(a.x) { return fabs(a.x) + fabs(a.y); }
__device__ inline bool nonzero(double a) { return a != 0; }
template <typename T> __device__ inline bool nonzero(T a) { return a.x != 0 || a.y != 0; }
template <typename T> __device__ inline T zero_v() { T r; r.x = 0; r.y = 0; return r; }
template <> __device__ inline double zero_v<double>() { return 0; }
template <typename T> __device__ inline T unit_v() { T r; r.x = 1; r.y = 0; return r; }
template <> __device__ inline double unit_v<double>() { return 1; }
// acc += a b, real and imaginary parts each as one expression
template <typename T> __device__ inline void mac(T &acc, T a, T b)
{
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}
__device__ inline void mac(double &acc, double a, double b) { acc += a * b; }
// acc += conj(a) b (conj = identity for real)
template <typename T> __device__ inline void macc(T &acc, T a, T b, bool cj)
{
  if (cj) a.y = -a.y;
  mac(acc, a, b);
}
__device__ inline void macc(double &acc, double a, double b, bool) { acc += a * b; }
template <typename T> __device__ inline T conj_if(T a, bool cj) { if (cj) a.y = -a.y; return a; }
__device__ inline double conj_if(double a, bool) { return a; }
template <typename T> __device__ inline double re_of(T a) { return a.x; }
template <typename T> __device__ inline double im_of(T a) { return a.y; }
__device__ inline double re_of(double a) { return a; }
__device__ inline double im_of(double) { return 0; }
template <typename T> struct IsReal { static const bool v = false; };
template <> struct IsReal<double> { static const bool v = true; };

template <typename T> __device__ inline T cdiv(T a, T b)
{
  typedef decltype(a.x) R;
  T r;
  if (fabs(b.x) >= fabs(b.y)) { R t = b.y / b.x, d = b.x + b.y * t; r.x = (a.x + a.y * t) / d; r.y = (a.y - a.x * t) / d; }
  else { R t = b.x / b.y, d = b.x * t + b.y; r.x = (a.x * t + a.y) / d; r.y = (a.y * t - a.x) / d; }
  return r;
}

// Strides between the instances of a batch of equal-size factorizations:
// matrix elements (a), pivot entries (p) and bytes of each instance's
// panel scratch (s).  A single factorization is a batch of one, strides 0;
// batched kernels take the instance from blockIdx.y or blockIdx.z.
struct BStride {
  long a = 0, p = 0, s = 0;
};
template <typename P> __device__ inline P *bshift(P *p, long bytes)
{
  return p ? (P *)((char *)p + bytes) : p;
}

// ---------------------------------------------------------------------
// C[m x n] -= A[m x k] B[k x n]  (ACC)  or  C = A B  (!ACC);
// 64x64 block tile, 4 waves of 32x32, K staged 16 at a time through LDS
// as separate real/imag planes, four real MFMAs per complex product.
// ---------------------------------------------------------------------
#define GBM 64
#define GBN 64
#define GBK 16

// G3M (complex64 only): three real products per complex product instead of
// four (Gauss / "3M": T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi),
// Re = T1 - T2, Im = T3 - T1 - T2; normwise stable, Higham, SIAM J. Matrix
// Anal. Appl. 13 (1992) 681-687), a quarter fewer MFMAs
template <typename T, bool ACC, bool G3M = false>
__global__ void __launch_bounds__(256)
k_gemm(long m, long n, long k, const T *A, long lda, const T *B, long ldb, T *C, long ldc, long bsa)
{
  {
    const long z = blockIdx.z;   // instance of a batch (bsa elements apart)
    A += z * bsa; B += z * bsa; C += z * bsa;
  }
  typedef typename CT<T>::R R;
  typedef typename CT<T>::V4 V4;
  __shared__ R sAr[GBK][GBM + 4], sAi[GBK][GBM + 4], sBr[GBK][GBN + 4], sBi[GBK][GBN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const long row0 = (long)blockIdx.x * GBM, col0 = (long)blockIdx.y * GBN;
  V4 accr[2][2], acci[2][2], acc3[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) { accr[a][b] = CT<T>::zero(); acci[a][b] = CT<T>::zero(); acc3[a][b] = CT<T>::zero(); }
  // tiles of chunk k0 into registers: all eight global loads in flight
  // before any LDS store
  auto load = [&](long k0, T (&v)[4], T (&u)[4]) {
    if (row0 + GBM <= m && col0 + GBN <= n && k0 + GBK <= k) {
#pragma unroll
      for (int s = 0; s < 4; s++) {
        v[s] = A[(row0 + (tid & 63)) + (k0 + (tid >> 6) + 4 * s) * lda];
        u[s] = B[(k0 + (tid & 15)) + (col0 + (tid >> 4) + 16 * s) * ldb];
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; s++) {
        int i = tid & 63, kk = (tid >> 6) + 4 * s;
        long gi = row0 + i, gk = k0 + kk;
        if (gi < m && gk < k) v[s] = A[gi + gk * lda]; else { v[s] = zero_v<T>(); }
        int kb = tid & 15, j = (tid >> 4) + 16 * s;
        long gkb = k0 + kb, gj = col0 + j;
        if (gkb < k && gj < n) u[s] = B[gkb + gj * ldb]; else { u[s] = zero_v<T>(); }
      }
    }
  };
  // the next chunk's loads are issued before this chunk's products
  // (complex64: 98 registers, four waves per SIMD; complex128: two)
  constexpr bool PF = true;
  T v[4], u[4];
  if (PF && k > 0) load(0, v, u);
  for (long k0 = 0; k0 < k; k0 += GBK) {
    if (!PF) load(k0, v, u);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int i = tid & 63, kk = (tid >> 6) + 4 * s;
      sAr[kk][i] = v[s].x; sAi[kk][i] = v[s].y;
      int kb = tid & 15, j = (tid >> 4) + 16 * s;
      sBr[kb][j] = u[s].x; sBi[kb][j] = u[s].y;
    }
    __syncthreads();
    if (PF && k0 + GBK < k) load(k0 + GBK, v, u);
    for (int kk0 = 0; kk0 < GBK; kk0 += 4) {
      int kk = kk0 + (lane >> 4);
      R ar[2], ai[2], br[2], bi[2];
      for (int t = 0; t < 2; t++) {
        ar[t] = sAr[kk][wr * 32 + t * 16 + (lane & 15)];
        ai[t] = sAi[kk][wr * 32 + t * 16 + (lane & 15)];
        br[t] = sBr[kk][wc * 32 + t * 16 + (lane & 15)];
        bi[t] = sBi[kk][wc * 32 + t * 16 + (lane & 15)];
      }
      if constexpr (G3M) {
        // accr: T1, acci: T2, acc3: T3
        R as[2], bs[2];
        for (int t = 0; t < 2; t++) { as[t] = ar[t] + ai[t]; bs[t] = br[t] + bi[t]; }
        for (int ti = 0; ti < 2; ti++)
          for (int tj = 0; tj < 2; tj++) {
            accr[ti][tj] = CT<T>::mfma(ar[ti], br[tj], accr[ti][tj]);
            acci[ti][tj] = CT<T>::mfma(ai[ti], bi[tj], acci[ti][tj]);
            acc3[ti][tj] = CT<T>::mfma(as[ti], bs[tj], acc3[ti][tj]);
          }
      } else {
        for (int ti = 0; ti < 2; ti++)
          for (int tj = 0; tj < 2; tj++) {
            // acc = Re/Im of -(A B) when accumulating into C, +(A B) otherwise
            const R sg = ACC ? (R)-1 : (R)1;
            accr[ti][tj] = CT<T>::mfma(sg * ar[ti], br[tj], accr[ti][tj]);
            accr[ti][tj] = CT<T>::mfma(-sg * ai[ti], bi[tj], accr[ti][tj]);
            acci[ti][tj] = CT<T>::mfma(sg * ar[ti], bi[tj], acci[ti][tj]);
            acci[ti][tj] = CT<T>::mfma(sg * ai[ti], br[tj], acci[ti][tj]);
          }
      }
    }
    __syncthreads();
  }
  for (int ti = 0; ti < 2; ti++)
    for (int tj = 0; tj < 2; tj++)
      for (int r = 0; r < 4; r++) {
        long gi = row0 + wr * 32 + ti * 16 + CT<T>::row(lane, r);
        long gj = col0 + wc * 32 + tj * 16 + (lane & 15);
        if (gi < m && gj < n) {
          R re = accr[ti][tj][r], im = acci[ti][tj][r];
          if constexpr (G3M) {
            const R t1 = accr[ti][tj][r], t2 = acci[ti][tj][r];
            re = ACC ? t2 - t1 : t1 - t2;
            im = ACC ? (t1 + t2) - acc3[ti][tj][r] : acc3[ti][tj][r] - (t1 + t2);
          }
          T c;
          if (ACC) { c = C[gi + gj * ldc]; c.x += re; c.y += im; }
          else { c.x = re; c.y = im; }
          C[gi + gj * ldc] = c;
        }
      }
}

// Real (double) C -= A B or C = A B: the same 64x64 tile and k-chunk
// staging, one v_mfma_f64_16x16x4f64 per product; each wave owns a 32x64
// strip (2 x 4 tiles of 16x16), so one LDS read of A feeds four products.
template <bool ACC>
__global__ void __launch_bounds__(256)
k_gemm_r(long m, long n, long k, const double *A, long lda, const double *B, long ldb, double *C, long ldc, long bsa)
{
  {
    const long z = blockIdx.z;
    A += z * bsa; B += z * bsa; C += z * bsa;
  }
  __shared__ double sA[GBK][2 * GBM + 4], sB[GBK][GBN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long row0 = (long)blockIdx.x * 2 * GBM, col0 = (long)blockIdx.y * GBN;
  d4 acc[2][4];
  for (int a = 0; a < 2; a++) for (int b = 0; b < 4; b++) acc[a][b] = d4{0, 0, 0, 0};
  // A chunk: 128 rows x 16, B chunk: 16 x 64 per k-chunk
  auto load = [&](long k0, double (&v)[8], double (&u)[4]) {
    if (row0 + 2 * GBM <= m && col0 + GBN <= n && k0 + GBK <= k) {
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = A[(row0 + (tid & 127)) + (k0 + (tid >> 7) + 2 * q) * lda];
#pragma unroll
      for (int q = 0; q < 4; q++) u[q] = B[(k0 + (tid & 15)) + (col0 + (tid >> 4) + 16 * q) * ldb];
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        long gi = row0 + (tid & 127), gk = k0 + (tid >> 7) + 2 * q;
        v[q] = (gi < m && gk < k) ? A[gi + gk * lda] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        long gk = k0 + (tid & 15), gj = col0 + (tid >> 4) + 16 * q;
        u[q] = (gk < k && gj < n) ? B[gk + gj * ldb] : 0.0;
      }
    }
  };
  double v[8], u[4];
  if (k > 0) load(0, v, u);
  for (long k0 = 0; k0 < k; k0 += GBK) {
#pragma unroll
    for (int q = 0; q < 8; q++) sA[(tid >> 7) + 2 * q][tid & 127] = v[q];
#pragma unroll
    for (int q = 0; q < 4; q++) sB[tid & 15][(tid >> 4) + 16 * q] = u[q];
    __syncthreads();
    if (k0 + GBK < k) load(k0 + GBK, v, u);
    for (int kk0 = 0; kk0 < GBK; kk0 += 4) {
      const int kk = kk0 + (lane >> 4);
      double ar[2], br[4];
      for (int t = 0; t < 2; t++) ar[t] = sA[kk][wave * 32 + t * 16 + (lane & 15)];
      for (int t = 0; t < 4; t++) br[t] = sB[kk][t * 16 + (lane & 15)];
      const double sg = ACC ? -1.0 : 1.0;
      for (int ti = 0; ti < 2; ti++)
        for (int tj = 0; tj < 4; tj++)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * ar[ti], br[tj], acc[ti][tj], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int ti = 0; ti < 2; ti++)
    for (int tj = 0; tj < 4; tj++)
      for (int r = 0; r < 4; r++) {
        long gi = row0 + wave * 32 + ti * 16 + (lane >> 4) + 4 * r;
        long gj = col0 + tj * 16 + (lane & 15);
        if (gi < m && gj < n) {
          if (ACC) C[gi + gj * ldc] += acc[ti][tj][r];
          else C[gi + gj * ldc] = acc[ti][tj][r];
        }
      }
}

static int gemm_kind(const double2 *) { return PK_GEMM_C128; }
static int gemm_kind(const float2 *) { return PK_GEMM_C64; }
static int gemm_kind(const double *) { return PK_GEMM_F64; }

// lds_pad: extra (unused) LDS per workgroup, which caps how many of its
// workgroups share a CU
template <typename T, bool ACC>
static void gemm(long m, long n, long k, const T *A, long lda, const T *B, long ldb, T *C, long ldc,
                 hipStream_t s, unsigned lds_pad = 0, int nbat = 1, long bsa = 0, bool g3m = false)
{
  if (m <= 0 || n <= 0 || k <= 0 || nbat <= 0) return;
  int tok = prof_begin(s, gemm_kind(A));
  if constexpr (IsReal<T>::v) {
    dim3 grid((unsigned)((m + 2 * GBM - 1) / (2 * GBM)), (unsigned)((n + GBN - 1) / GBN), (unsigned)nbat);
    hipLaunchKernelGGL((k_gemm_r<ACC>), grid, dim3(256), lds_pad, s, m, n, k, A, lda, B, ldb, C, ldc, bsa);
    prof_end(s, tok, 2.0 * (double)m * (double)n * (double)k * nbat);
  } else {
    dim3 grid((unsigned)((m + GBM - 1) / GBM), (unsigned)((n + GBN - 1) / GBN), (unsigned)nbat);
    if (std::is_same<T, float2>::value && ACC && g3m)
      hipLaunchKernelGGL((k_gemm<T, ACC, true>), grid, dim3(256), lds_pad, s, m, n, k, A, lda, B, ldb, C, ldc, bsa);
    else
      hipLaunchKernelGGL((k_gemm<T, ACC>), grid, dim3(256), lds_pad, s, m, n, k, A, lda, B, ldb, C, ldc, bsa);
    prof_end(s, tok, 8.0 * (double)m * (double)n * (double)k * nbat);
  }
}

template <typename T>
static void gemm_sub(long m, long n, long k, const T *A, long lda, const T *B, long ldb, T *C, long ldc,
                     hipStream_t s, unsigned lds_pad = 0, int nbat = 1, long bsa = 0, bool g3m = false)
{
  gemm<T, true>(m, n, k, A, lda, B, ldb, C, ldc, s, lds_pad, nbat, bsa, g3m);
}

// The mixed solver's complex64 factorization (a preconditioner for the
// complex128 refinement) updates with three real products per complex
// product; SCUFFX_GEMM_3M=0, read at every factorization, keeps four (its
// normwise error is 2-4x smaller, which matters only past condition
// numbers of ~1e7, where complex64 factors barely precondition anyway)
static bool mixed_gemm_3m()
{
  const char *e = getenv("SCUFFX_GEMM_3M");
  return !(e && !strcmp(e, "0"));
}

// ---------------------------------------------------------------------
// cooperative panel factorization: rows [r0, r0+m), columns [r0, r0+w),
// w <= PW.  Workgroup g keeps rows [r0 + g*rows, ...) of the panel in
// LDS for the whole panel.  Per column one grid barrier: before it every
// workgroup publishes its best |re|+|im| candidate together with that
// candidate's row, and the owner of the diagonal row publishes that row;
// after it every workgroup picks the global pivot (largest, first row on
// ties, as izamax) and has both rows of the interchange at hand.  The
// published slots alternate between two buffers so that a workgroup
// running ahead never overwrites a slot another one still reads.
// ---------------------------------------------------------------------
#define PW 32

template <typename T> struct PanelRows;
template <> struct PanelRows<double2> { static const int n = 128; };
template <> struct PanelRows<float2> { static const int n = 256; };
template <> struct PanelRows<double> { static const int n = 256; };

struct PanelScratch {
  unsigned *bar;   // grid barrier: [0] arrivals, [32] generation, [64] timeout flag (own cache lines)
  double *candv;   // [2][Gmax]
  unsigned long long *ckey;   // [2][Gmax]: (local candidate row << 32) | column sequence number
  void *candrow;   // [2][Gmax][PW] T
  void *rowc;      // [2][PW] T
  int *plan;       // [2*PW+1]: dst rows, src rows, count of the panel's net row permutation
  void *linv;      // [npanels][PW][PW] T or null: inverse of each panel's unit-lower block (column-major)
  unsigned long long *ts;   // tournament-pivoting sets and root area (k_panel_ts), or null
  // the same scratch of the instance `bytes` further on (a batch's scratch
  // areas are equal and contiguous)
  __device__ inline PanelScratch shift(long bytes) const
  {
    PanelScratch o = *this;
    o.bar = bshift(bar, bytes); o.candv = bshift(candv, bytes); o.ckey = bshift(ckey, bytes);
    o.candrow = bshift((char *)candrow, bytes); o.rowc = bshift((char *)rowc, bytes);
    o.plan = bshift(plan, bytes); o.linv = bshift((char *)linv, bytes); o.ts = bshift(ts, bytes);
    return o;
  }
};


// Cross-workgroup exchange inside the panel without grid barriers: the
// shared words are read and written only by agent-scope relaxed atomics
// (coherent across the XCDs' L2s), and a writer waits for its data
// stores to complete (s_waitcnt) before it stores the sequence-tagged key
// that readers poll.  No release/acquire fences, hence no L2 write-back or
// invalidation while the trailing GEMM runs beside the panel.
__device__ inline void st_coh(unsigned long long *p, unsigned long long v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_coh(const unsigned long long *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coh(float2 *p, float2 v)
{
  unsigned long long u = ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
  st_coh((unsigned long long *)p, u);
}
__device__ inline float2 ld_coh(const float2 *p)
{
  unsigned long long u = ld_coh((const unsigned long long *)p);
  float2 v;
  v.x = __uint_as_float((unsigned)u);
  v.y = __uint_as_float((unsigned)(u >> 32));
  return v;
}
__device__ inline void st_coh(double2 *p, double2 v)
{
  st_coh((unsigned long long *)p, (unsigned long long)__double_as_longlong(v.x));
  st_coh((unsigned long long *)p + 1, (unsigned long long)__double_as_longlong(v.y));
}
__device__ inline void st_coh(double *p, double v)
{
  st_coh((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
}
__device__ inline double ld_coh(const double *p)
{
  return __longlong_as_double((long long)ld_coh((const unsigned long long *)p));
}
__device__ inline double2 ld_coh(const double2 *p)
{
  double2 v;
  v.x = __longlong_as_double((long long)ld_coh((const unsigned long long *)p));
  v.y = __longlong_as_double((long long)ld_coh((const unsigned long long *)p + 1));
  return v;
}
// polls before a wait gives up (about a second); the kernel then
// finishes and the host reports the time-out
#define GBAR_SPIN (1u << 20)
// keys one lane of the polling wave holds: up to 64 * 8 workgroups
#define KEYS_PER_LANE 8

// every memory operation this thread issued has completed; the signal
// fences keep the compiler from moving memory operations across it
__device__ inline void drain()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <typename T>
__global__ void __launch_bounds__(PanelRows<T>::n)
k_panel(T *A, long lda, long r0, long m, int w, int *ipiv, int *info, PanelScratch S, int Gmax, unsigned seq0)
{
  typedef typename CT<T>::R R;
  constexpr int NR = PanelRows<T>::n;
  constexpr bool kPackKey = std::is_same<T, float2>::value;
  __shared__ T slab[PW][NR];           // column-major: slab[q][t] = A(lo + t, r0 + q)
  __shared__ R wv[NR / 64];
  __shared__ int wi[NR / 64];
  __shared__ T spiv[PW], sold[PW];
  __shared__ long spv[PW];
  __shared__ int s_p;
  __shared__ R s_bv;
  __builtin_amdgcn_s_setprio(3);   // critical path beside the trailing GEMM
  const int t = threadIdx.x, G = gridDim.x, g = blockIdx.x, lane = t & 63, wave = t >> 6;
  const long lo = r0 + (long)g * NR;
  const int nrow = (int)std::min<long>(NR, r0 + m - lo);
  const long i = lo + t;                 // this thread's row
  T *candrow = (T *)S.candrow, *rowc = (T *)S.rowc;
  unsigned *err = S.bar + 64;
  __shared__ int s_gw;
  if (t < nrow) {
    T v[PW];
    if (w == PW) {
#pragma unroll
      for (int q = 0; q < PW; q++) v[q] = A[i + (r0 + q) * lda];
#pragma unroll
      for (int q = 0; q < PW; q++) slab[q][t] = v[q];
    } else {
      for (int q = 0; q < w; q++) slab[q][t] = A[i + (r0 + q) * lda];
    }
  }
  __syncthreads();
  for (int j = 0; j < w; j++) {
    const long c = r0 + j;
    const int buf = j & 1;
    const unsigned seq = seq0 + (unsigned)j;
    // local candidate: largest |re|+|im| among own rows >= c, first on ties
    R best = -1; int bi = -1;
    if (t < nrow && i >= c) { T v = slab[j][t]; best = mag1(v); bi = t; }
    for (int off = 32; off > 0; off >>= 1) {
      R ob = __shfl_xor(best, off);
      int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi >= 0 && (bi < 0 || oi < bi))) { best = ob; bi = oi; }
    }
    if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
    __syncthreads();
    R b0 = wv[0]; int bb = wi[0];
    for (int q = 1; q < NR / 64; q++)
      if (wv[q] > b0 || (wv[q] == b0 && wi[q] >= 0 && (bb < 0 || wi[q] < bb))) { b0 = wv[q]; bb = wi[q]; }
    // publish: candidate row (and row c, by its owner), then value, then key
    if (t < w) {
      if (bb >= 0) st_coh(candrow + ((long)buf * Gmax + g) * PW + t, slab[t][bb]);
      if (c >= lo && c < lo + nrow) st_coh(rowc + buf * PW + t, slab[t][c - lo]);
    }
    // complex64: the value travels in the key itself (value bits << 32 |
    // local row << 23 | 23-bit sequence number, row 511 = no candidate), so
    // the winner is known from the polled keys alone; complex128 needs all
    // 64 bits of its value and publishes it beside the key
    if (!kPackKey && t == 0)
      st_coh((unsigned long long *)S.candv + buf * Gmax + g, (unsigned long long)__double_as_longlong((double)b0));
    drain();
    __syncthreads();
    if (t == 0) {
      unsigned long long key;
      if (kPackKey)
        key = ((unsigned long long)__float_as_uint((float)b0) << 32) |
              ((unsigned long long)(bb >= 0 ? bb : 511) << 23) | (seq & 0x7FFFFFu);
      else
        key = ((unsigned long long)(unsigned)(bb >= 0 ? bb : -1) << 32) | seq;
      st_coh(S.ckey + buf * Gmax + g, key);
    }
    // global pivot: wave 0 waits for every workgroup's key of this column
    if (t < 64) {
      // bounded: past GBAR_SPIN polls the error word is set, and once it
      // is set no workgroup waits again (the panel then finishes with
      // meaningless values and the host reports the time-out)
      // the keys of the last poll stay in registers (G <= 64 * KEYS_PER_LANE)
      unsigned long long kk[KEYS_PER_LANE];
      const unsigned smask = kPackKey ? 0x7FFFFFu : 0xFFFFFFFFu;
      unsigned spins = 0;
      for (;;) {
        bool ok = true;
#pragma unroll
        for (int u = 0; u < KEYS_PER_LANE; u++) {
          const int q = lane + 64 * u;
          if (q < G) {
            kk[u] = ld_coh(S.ckey + buf * Gmax + q);
            if (((unsigned)kk[u] & smask) != (seq & smask)) ok = false;
          }
        }
        if (__ballot(!ok) == 0) break;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (++spins > GBAR_SPIN) { __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
        __builtin_amdgcn_s_sleep(1);
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      R b = -1; long p = -1; int gw = -1;
#pragma unroll
      for (int u = 0; u < KEYS_PER_LANE; u++) {
        const int q = lane + 64 * u;
        if (q >= G) continue;
        const unsigned long long k = kk[u];
        int li;
        R v;
        if (kPackKey) {
          li = (int)((k >> 23) & 511u);
          if (li == 511) continue;
          v = (R)__uint_as_float((unsigned)(k >> 32));
        } else {
          li = (int)(unsigned)(k >> 32);
          if (li < 0) continue;
          v = (R)__longlong_as_double((long long)ld_coh((const unsigned long long *)S.candv + buf * Gmax + q));
        }
        long ii = r0 + (long)q * NR + li;
        if (v > b || (v == b && ii < p)) { b = v; p = ii; gw = q; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        R ob = __shfl_xor(b, off);
        long op = __shfl_xor(p, off);
        int og = __shfl_xor(gw, off);
        if (op >= 0 && (ob > b || (ob == b && (p < 0 || op < p)))) { b = ob; p = op; gw = og; }
      }
      if (t == 0) { s_p = p < 0 ? (int)c : (int)p; s_gw = gw; s_bv = b; }
    }
    __syncthreads();
    const long p = s_p;
    if (t < w) {
      sold[t] = ld_coh(rowc + buf * PW + t);
      spiv[t] = s_gw >= 0 ? ld_coh(candrow + ((long)buf * Gmax + s_gw) * PW + t) : sold[t];
    }
    if (t == 0) spv[j] = p;
    if (g == 0 && t == 0) {
      ipiv[c] = (int)(p + 1);
      if (s_bv == 0 && *info == 0) *info = (int)(c + 1);
    }
    __syncthreads();
    if (t < w) {
      if (c >= lo && c < lo + nrow) slab[t][c - lo] = spiv[t];
      if (p != c && p >= lo && p < lo + nrow) slab[t][p - lo] = sold[t];
    }
    __syncthreads();
    T piv = spiv[j];
    const bool nz = nonzero(piv);
    if (t < nrow && i > c) {
      T l = slab[j][t];
      if (nz) { l = cmul(l, crecip(piv)); slab[j][t] = l; }
      for (int q = j + 1; q < w; q++) slab[q][t] = csub(slab[q][t], cmul(l, spiv[q]));
    }
    __syncthreads();
  }
  for (int q = 0; q < w; q++)
    if (t < nrow) A[i + (r0 + q) * lda] = slab[q][t];
  // the inverse of the panel's unit-lower w x w block (workgroup 0 holds
  // its rows), column t by substitution, identity outside w x w
  if (S.linv && g == 0 && t < PW) {
    T x[PW];
#pragma unroll
    for (int r = 0; r < PW; r++) {
      T v = (r == t) ? unit_v<T>() : zero_v<T>();
      if (r < w && t < w) {
#pragma unroll
        for (int k = 0; k < r; k++) v = csub(v, cmul(slab[k][r], x[k]));
      }
      x[r] = v;
    }
    T *li = (T *)S.linv + (r0 / PW) * PW * PW + (long)t * PW;
#pragma unroll
    for (int r = 0; r < PW; r++) li[r] = x[r];
  }
  // net permutation of the panel's w interchanges on the rows they touch,
  // one row per lane of wave 0: lanes 0..w-1 are rows r0.., pivot rows
  // from below are appended.  After the interchanges row dst[q] holds
  // what row src[q] held before.
  if (g == 0 && t < 64) {
    long row = (t < w) ? r0 + t : -1;
    int content = t, cnt = w;
    for (int j = 0; j < w; j++) {
      const long pj = spv[j];
      if (pj == r0 + j) continue;
      int ip;
      if (pj < r0 + w) ip = (int)(pj - r0);
      else {
        unsigned long long hit = __ballot(row == pj);
        if (hit) ip = __ffsll(hit) - 1;
        else { ip = cnt; if (t == cnt) { row = pj; content = cnt; } cnt++; }
      }
      int cj = __shfl(content, j), cp = __shfl(content, ip);
      if (t == j) content = cp;
      if (t == ip) content = cj;
    }
    long srow = __shfl(row, content);
    if (t < cnt) { S.plan[t] = (int)row; S.plan[PW * 2 + t] = (int)srow; }
    if (t == 0) S.plan[PW * 4] = cnt;
  }
}

// ---------------------------------------------------------------------
// Tournament-pivoting panel for the complex64 factorization of the mixed
// solver (communication-avoiding LU, "TSLU": Grigori, Demmel and Xiang,
// SIAM J. Matrix Anal. Appl. 32 (2011) 1317-1350).  Each workgroup factors
// its 256 rows x w columns in LDS with partial pivoting and nominates the w
// rows it pivoted on; groups of TS_FAN nominee sets are stacked and factored
// again, up a tree, until one workgroup factors the last stack: its w pivot
// rows, in order, are the panel's pivots and its factors the w x w diagonal
// block L11\U11.  Every workgroup then writes its rows of the permuted panel:
// the diagonal block, or L21 rows = (original row) U11^-1 by substitution.
// Pivot choices differ from row-by-row partial pivoting (the complex128
// path keeps that); the mixed solver's refinement in complex128 takes the
// factors as a preconditioner.  Cross-workgroup data moves only through
// agent-scope relaxed atomics with drain-then-flag, as in k_panel: three
// waits per panel (two at most 64 workgroups) instead of one per column.
// ---------------------------------------------------------------------
#define TS_FAN 8
#define TS_SET 128                  // u64 words of one nominee set: row indices [PW], flag,
#define TS_IDX 0                    // references [PW] (row of each nominee's original values in
#define TS_FLAG (PW + 8)            // the copy area: workgroup * 256 + row)
#define TS_REF (PW + 16)
#define TS_TOP (PW * PW)            // root area: final block [PW][PW], (spare) [PW][PW], plan, flag
#define TS_PLAN (2 * PW * PW)
#define TS_NPLAN (2 * PW * PW + 2 * PW)
#define TS_FIN (2 * PW * PW + 2 * PW + 8)
#define TS_UREP 2176                // first replica (workgroup g reads replica g % 8)
#define TS_REP 1056                 // u64 words between replicas of the final block
#define TS_ROOT (2176 + 8 * TS_REP) // root area + 8 replicas of the final block

// the tournament's scratch: sets of all levels [levels + 1][Gmax], the
// copy area [Gmax][256][PW] (each workgroup's rows as loaded), the root area
__host__ __device__ inline long ts_copy_off(int levels, long Gmax) { return (long)(levels + 1) * Gmax * TS_SET; }
__host__ __device__ inline long ts_root_off(int levels, long Gmax) { return ts_copy_off(levels, Gmax) + Gmax * 256 * PW; }

__host__ __device__ inline int ts_levels(long G)
{
  int l = 0;
  for (long c = G; c > 1; c = (c + TS_FAN - 1) / TS_FAN) l++;
  return l;
}

// 1/a for complex64 with hardware reciprocals (about 1 ulp each; the
// tournament panel's pivots), scaled by max(|re|, |im|) against overflow
// and underflow of |a|^2
__device__ inline float2 crecip_fast(float2 a)
{
  const float s = fmaxf(fabsf(a.x), fabsf(a.y)), is = __builtin_amdgcn_rcpf(s);
  const float xs = a.x * is, ys = a.y * is;
  const float d = __builtin_amdgcn_rcpf(fmaf(xs, xs, ys * ys)) * is;
  float2 r;
  r.x = xs * d;
  r.y = -ys * d;
  return r;
}

// r - l c with fused multiply-adds (complex64)
__device__ inline float2 cfms(float2 r, float2 l, float2 c)
{
  float2 o;
  o.x = fmaf(-l.x, c.x, fmaf(l.y, c.y, r.x));
  o.y = fmaf(-l.x, c.y, fmaf(-l.y, c.x, r.y));
  return o;
}

// maximum of a 32-bit key over the wave, in every lane: within each row
// of 16 lanes by DPP (quad swaps, half-row and row mirrors), then over the
// four rows by lane reads (the result is wave-uniform).  All lanes active.
__device__ inline unsigned wave_max_key(unsigned k)
{
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x141, 0xF, 0xF, false));   // row_half_mirror
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x140, 0xF, 0xF, false));   // row_mirror
  const unsigned a = __builtin_amdgcn_readlane(k, 0), b = __builtin_amdgcn_readlane(k, 16);
  const unsigned c = __builtin_amdgcn_readlane(k, 32), d = __builtin_amdgcn_readlane(k, 48);
  return max(max(a, b), max(c, d));
}

// partial-pivoting LU (largest |re|+|im|, lowest row on ties) of the rows
// [0, nr) the threads of the workgroup hold in registers (thread t: row t,
// r[0..w-1]), rows left in place: spiv[j] = row of pivot j, srank[t] = j
// on pivot rows (else -1); each row ends as its row of the packed factors
// (multipliers left of its pivot step, or its U row).  Returns min(w, nr);
// *zero = first step with a zero pivot or -1.  The pivot search compares
// keys bits(|re|+|im|) + 1 (0: row already a pivot or absent), which order
// as the non-negative magnitudes do.  One barrier per column: the winning
// row of each wave goes through LDS (two buffers: a wave writes buffer
// j & 1 again only after every wave has passed step j + 1's barrier, i.e.
// has read it).
__device__ __attribute__((always_inline)) inline int ts_lu(float2 (&r)[PW], int nr, int w, int *spiv, int *srank,
                                                           unsigned long long (*wkv)[4], float2 (*cand)[4][PW],
                                                           int *zero)
{
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  bool used = t >= nr;
  srank[t] = -1;
  const int np = min(w, nr);
  int z = -1;
#pragma unroll
  for (int j = 0; j < PW; j++) {
    if (j < np) {
      const unsigned key = used ? 0u : __float_as_uint(fabsf(r[j].x) + fabsf(r[j].y)) + 1u;
      const unsigned wm = wave_max_key(key);
      const unsigned long long hit = __ballot(wm != 0 && key == wm);
      const int wl = hit ? __ffsll(hit) - 1 : -1;
      if (lane == wl) {
#pragma unroll
        for (int q = j; q < PW; q++) cand[j & 1][wave][q] = r[q];
        // the wave's key and row in one word: the largest word has the
        // largest key and, among equal keys, the lowest row
        wkv[j & 1][wave] = ((unsigned long long)wm << 32) | (unsigned)(255 - t);
      }
      if (lane == 0 && wl < 0) wkv[j & 1][wave] = 0ull;
      __syncthreads();
      unsigned long long best = wkv[j & 1][0];
#pragma unroll
      for (int q = 1; q < 4; q++) best = max(best, wkv[j & 1][q]);
      const unsigned b0 = (unsigned)(best >> 32);
      const int p = 255 - (int)(unsigned)best;
      const int pw = p >> 6;
      if (t == p) { used = true; srank[t] = j; }
      if (t == 0) spiv[j] = p;
      if (b0 == 1u && z < 0) z = j;
      if (!used) {
        float2 l = r[j];
        if (b0 != 1u) { l = cmul(l, crecip_fast(cand[j & 1][pw][j])); r[j] = l; }
#pragma unroll
        for (int q = j + 1; q < PW; q++) r[q] = cfms(r[q], l, cand[j & 1][pw][q]);
      }
    }
  }
  __syncthreads();
  *zero = z;
  return np;
}

// nominees in the level-l set of workgroup g: min(w, rows of workgroups g
// .. g + TS_FAN^l - 1), by induction over the merges
__device__ inline int ts_count(long g, long span, long m, int w)
{
  const long rows = min(span * 256, m - g * 256);
  return (int)min((long)w, rows);
}

// wave 0 waits until the n <= 64 words f[k * stride] hold seq (bounded as
// in k_panel), then the workgroup synchronizes
__device__ void ts_wait(const unsigned long long *f, long stride, int n, unsigned long long seq, unsigned *err)
{
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    unsigned spins = 0;
    for (;;) {
      bool ok = true;
      if (lane < n) ok = ld_coh(f + lane * stride) == seq;
      if (__ballot(!ok) == 0) break;
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (++spins > GBAR_SPIN) { __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
      __builtin_amdgcn_s_sleep(1);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
  __syncthreads();
}

// panel rows [r0, r0 + m) x columns [r0, r0 + w), w <= PW; grid: one
// workgroup per 256 rows, all co-resident; outputs as k_panel (factors in
// place, ipiv, info, the net row permutation in S.plan, the inverse of the
// unit-lower diagonal block in S.linv)
// (at most 168 VGPRs, so that a workgroup fits on SIMDs that run three
// waves of the trailing GEMM instead of waiting for them to retire)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_panel_ts(float2 *A, long lda, long r0, long m, int w, int *ipiv, int *info, PanelScratch S, int Gmax, unsigned seq0,
           BStride bs)
{
  constexpr int NR = 256;
  __shared__ float2 sF[PW][PW + 1];     // final block, row-major
  __shared__ float2 cand[2][NR / 64][PW];
  __shared__ float2 sRd[PW];
  __shared__ unsigned long long wkv[2][NR / 64];
  __shared__ int spiv[PW], srank[NR], ssrc[NR], sfrom[NR];
  __shared__ long sidx[NR];
  // the panel is the critical path and the trailing GEMM shares its CUs:
  // its waves win instruction issue on their SIMDs
  {
    // instance blockIdx.y of a batch
    const long bi = blockIdx.y;
    A += bi * bs.a;
    ipiv += bi * bs.p;
    info = bshift(info, bi * bs.s);
    S = S.shift(bi * bs.s);
  }
  __builtin_amdgcn_s_setprio(3);
  // workgroups 0 .. G - 1 hold the panel's rows; workgroup G only inverts
  // the final block's unit-lower triangle, beside the others' last phase
  const int t = threadIdx.x, G = gridDim.x - 1, g = blockIdx.x;
  const long lo = r0 + (long)g * NR;
  const int nrow = (int)max(0L, min((long)NR, r0 + m - lo));
  const long i = lo + t;
  const unsigned long long seq = seq0;
  unsigned *err = S.bar + 64;
  unsigned long long *ts = S.ts;
  unsigned long long *copy = ts + ts_copy_off(ts_levels(Gmax), Gmax);
  unsigned long long *root = ts + ts_root_off(ts_levels(Gmax), Gmax);
  const int nl = ts_levels(G);
  if (g == G) {
    ts_wait(root + TS_FIN, 0, 1, seq, err);
    // (all of this thread's loads in flight before the first LDS store)
    float2 fv[PW * PW / NR];
#pragma unroll
    for (int u = 0; u < PW * PW / NR; u++) fv[u] = ld_coh((const float2 *)(root + t + u * NR));
#pragma unroll
    for (int u = 0; u < PW * PW / NR; u++) {
      const int e = t + u * NR, rr = e / PW, q = e % PW;
      if (rr < w && q < w) sF[rr][q] = fv[u];
    }
    __syncthreads();
    // the inverse of the unit-lower w x w block, column t by substitution
    if (S.linv && t < PW) {
      float2 x[PW];
#pragma unroll
      for (int rr = 0; rr < PW; rr++) {
        float2 val;
        val.x = (rr == t) ? 1 : 0; val.y = 0;
        if (rr < w && t < w) {
#pragma unroll
          for (int k = 0; k < rr; k++) val = cfms(val, sF[rr][k], x[k]);
        }
        x[rr] = val;
      }
      float2 *li = (float2 *)S.linv + (r0 / PW) * PW * PW + (long)t * PW;
#pragma unroll
      for (int rr = 0; rr < PW; rr++) li[rr] = x[rr];
    }
    return;
  }

  // own rows into registers (A keeps the original values until the last
  // phase, which re-reads them); workgroup 0 also publishes rows r0 .. r0 +
  // w - 1 (those the interchanges displace)
  // (columns from w on hold zeros; their loads repeat column w - 1, so
  // that the loads are not split by per-column branches)
  float2 r[PW];
#pragma unroll
  for (int q = 0; q < PW; q++) r[q] = zero_v<float2>();
  if (t < nrow) {
#pragma unroll
    for (int q = 0; q < PW; q++) {
      const float2 v = A[i + (r0 + min(q, w - 1)) * lda];
      r[q] = q < w ? v : zero_v<float2>();
    }
  }
  // every row's original values to the copy area (read by the merges and,
  // for the rows the interchanges displace from the diagonal block, by the
  // last phase); drained before this workgroup's first flag
  if (t < nrow) {
#pragma unroll
    for (int q = 0; q < PW; q++) st_coh((float2 *)(copy + ((long)g * NR + t) * PW + q), r[q]);
  }
  sidx[t] = lo + t;
  // stage 0: this workgroup's rows; stage l >= 1: workgroups g % TS_FAN^l
  // == 0 stack the nominee sets of g, g + sp, .., g + (TS_FAN - 1) sp (sp =
  // TS_FAN^(l-1)) from stage l - 1.  The last stage runs on workgroup 0.
  int zero = -1;
  long sp = 1;
  for (int l = 0; l <= nl; l++) {
    int nr = nrow;
    const unsigned long long *prev = ts + (long)(l > 0 ? l - 1 : 0) * Gmax * TS_SET;
    if (l > 0) {
      if (g % (sp * TS_FAN)) break;
      const int nch = (int)min((long)TS_FAN, (G - g + sp - 1) / sp);
      ts_wait(prev + (long)g * TS_SET + TS_FLAG, sp * TS_SET, nch, seq, err);
      // the stack: child k's nominees at rows sbase[k] ..
      int sbase[TS_FAN + 1];
      sbase[0] = 0;
#pragma unroll
      for (int k = 0; k < TS_FAN; k++)
        sbase[k + 1] = sbase[k] + (k < nch ? ts_count(g + k * sp, sp, m, w) : 0);
      nr = sbase[TS_FAN];
#pragma unroll
      for (int q = 0; q < PW; q++) { r[q] = zero_v<float2>(); }
      if (t < nr) {
        int k = 0;
#pragma unroll
        for (int q = 1; q < TS_FAN; q++)
          if (t >= sbase[q]) k = q;
        const int rr = t - sbase[k];
        const unsigned long long *cs = prev + (g + k * sp) * TS_SET;
        const int ref = (int)ld_coh(cs + TS_REF + rr);
        const unsigned long long *os = copy + (long)ref * PW;   // the row's original values
#pragma unroll
        for (int q = 0; q < PW; q++) r[q] = ld_coh((const float2 *)(os + q));
        sidx[t] = (long)ld_coh(cs + TS_IDX + rr);
        ssrc[t] = ref;
      }
    }
    const int np = ts_lu(r, nr, w, spiv, srank, wkv, cand, &zero);
    if (l == nl) {
      // workgroup 0: the final block in pivot order
      const int k = t < nr ? srank[t] : -1;
      if (k >= 0) {
#pragma unroll
        for (int q = 0; q < PW; q++) sF[k][q] = r[q];
      }
      break;
    }
    // nominees: original values of the pivot rows
    unsigned long long *set = ts + ((long)l * Gmax + g) * TS_SET;
    const int k = t < nr ? srank[t] : -1;
    if (k >= 0) {
      // references to the original values in the copy area
      st_coh(set + TS_REF + k, (unsigned long long)(l == 0 ? g * NR + t : ssrc[t]));
      st_coh(set + TS_IDX + k, (unsigned long long)sidx[t]);
    }
    (void)np;
    drain();
    __syncthreads();
    if (t == 0) st_coh(set + TS_FLAG, seq);
    if (l > 0) sp *= TS_FAN;
  }

  if (g == 0) {
    // publish the final block, and its upper triangle U11 in eight
    // replicas (read by all workgroups at once); ipiv, the net
    // permutation and info
    __syncthreads();
    for (int e = t; e < PW * PW; e += NR) {
      const int rr = e / PW, q = e % PW;
      if (rr < w && q < w) {
        st_coh((float2 *)(root + rr * PW + q), sF[rr][q]);
        if (q >= rr)
          for (int c = 0; c < 8; c++) st_coh((float2 *)(root + TS_UREP + c * TS_REP + rr * PW + q), sF[rr][q]);
      }
    }
    if (t < 64) {
      // lanes hold (position, original row now there) for every touched
      // position: lanes 0..w-1 the diagonal block's, pivot rows from
      // below appended (lane indices are wave-uniform: lane reads)
      int row = (t < w) ? (int)(r0 + t) : -1, content = row;
      const int Rl = (t < w) ? (int)sidx[spiv[t]] : -1;    // original row of pivot t
      int cnt = w;
      for (int j = 0; j < w; j++) {
        const int R = __builtin_amdgcn_readlane(Rl, j);
        const unsigned long long hit = __ballot(content == R);
        int ip;
        if (hit) ip = __ffsll(hit) - 1;
        else { ip = cnt; if (t == cnt) { row = R; content = R; } cnt++; }
        const int pos = __builtin_amdgcn_readlane(row, ip);
        if (t == 0) ipiv[r0 + j] = pos + 1;
        const int cj = __builtin_amdgcn_readlane(content, j), cp = __builtin_amdgcn_readlane(content, ip);
        if (t == j) content = cp;
        if (t == ip) content = cj;
      }
      if (t < cnt) {
        S.plan[t] = (int)row;
        S.plan[PW * 2 + t] = (int)content;
        st_coh(root + TS_PLAN + t, ((unsigned long long)row << 32) | (unsigned long long)content);
      }
      if (t == 0) {
        S.plan[PW * 4] = cnt;
        st_coh(root + TS_NPLAN, (unsigned long long)cnt);
        if (zero >= 0 && *info == 0) *info = (int)(r0 + zero + 1);
      }
    }
    drain();
    __syncthreads();
    if (t == 0) st_coh(root + TS_FIN, seq);
  }

  // every workgroup: its rows of the permuted panel.  The row's original
  // values (the panel never writes them before this phase, and only this
  // thread writes its row after it) are loaded before the root's flag is
  // awaited; a row displaced from the diagonal block reloads below.
  float2 v[PW];
#pragma unroll
  for (int q = 0; q < PW; q++) { v[q] = zero_v<float2>(); }
  if (t < nrow && i >= r0 + w) {
    // plain loads (opaque address, as above); columns from w on repeat
    // column w - 1 and are never used
    float2 *Ai = A + i + r0 * lda;
    asm volatile("" : "+v"(Ai));
#pragma unroll
    for (int q = 0; q < PW; q++) v[q] = Ai[min(q, w - 1) * lda];
  }
  ts_wait(root + TS_FIN, 0, 1, seq, err);
  // workgroup 0 (the diagonal block's rows) reads the whole final
  // block, the others U11 from replica g % 8; with the plan words, all
  // loads in flight at once
  float2 fv[PW * PW / NR];
  {
    const unsigned long long *fb = g == 0 ? root : root + TS_UREP + (g & 7) * TS_REP;
#pragma unroll
    for (int u = 0; u < PW * PW / NR; u++) fv[u] = ld_coh((const float2 *)(fb + t + u * NR));
  }
  int npl = 0;
  unsigned long long pe = 0;
  if (t < 2 * PW) {
    npl = (int)ld_coh(root + TS_NPLAN);
    pe = ld_coh(root + TS_PLAN + t);
  }
#pragma unroll
  for (int u = 0; u < PW * PW / NR; u++) {
    const int e = t + u * NR, rr = e / PW, q = e % PW;
    if (rr < w && q < w && (g == 0 || q >= rr)) sF[rr][q] = fv[u];
  }
  // displaced rows of this workgroup: sfrom[row - lo] = original row
  sfrom[t] = -1;
  __syncthreads();
  if (t < 2 * PW) {
    const long dst = (long)(pe >> 32);
    if (t < npl && dst >= lo && dst < lo + nrow) sfrom[dst - lo] = (int)(pe & 0xFFFFFFFFull);
  }
  __syncthreads();
  if (t < w) {
    const float2 d = sF[t][t];
    sRd[t] = (d.x != 0 || d.y != 0) ? crecip(d) : d;
  }
  __syncthreads();
  if (t < nrow) {
    if (i < r0 + w) {
#pragma unroll
      for (int q = 0; q < PW; q++)
        if (q < w) A[i + (r0 + q) * lda] = sF[i - r0][q];
    } else {
      const long c = sfrom[t] >= 0 ? sfrom[t] : i;
      if (c == i) {
        // the values loaded above
      } else if (c >= r0 && c < r0 + w) {
        // a row displaced from the diagonal block (no other source exists;
        // anything else is reported as a failed panel)
#pragma unroll
        for (int q = 0; q < PW; q++) v[q] = ld_coh((const float2 *)(copy + (c - r0) * PW + q));
      } else {
        __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // x U11 = (original row); a zero pivot leaves its column unscaled
#pragma unroll
      for (int q = 0; q < PW; q++) {
        if (q < w) {
          float2 x = v[q];
#pragma unroll
          for (int k = 0; k < q; k++) x = cfms(x, v[k], sF[k][q]);
          const float2 rd = sRd[q];
          if (rd.x != 0 || rd.y != 0) x = cmul(x, rd);
          v[q] = x;
        }
      }
#pragma unroll
      for (int q = 0; q < PW; q++)
        if (q < w) A[i + (r0 + q) * lda] = v[q];
    }
  }
}

// applies nplans consecutive panel row permutations (each: dst <- src on
// at most 2 PW rows, stride 4 PW + 1 ints) to the columns [c0, c1) except
// [s0, s1); one wave per column, four columns in flight per wave, plans in
// order within a column
#define PLAN_LEN (4 * PW + 1)
template <typename T>
__global__ void __launch_bounds__(256)
k_laswp_plans(T *A, long lda, long c0, long c1, long s0, long s1, const int *plans, int nplans, BStride bs)
{
  {
    const long bi = blockIdx.y;   // instance of a batch
    A += bi * bs.a;
    plans = bshift(plans, bi * bs.s);
  }
  const int lane = threadIdx.x & 63;
  const long skip = s1 - s0;
  const long ncol = (c1 - c0) - skip;
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  for (long c4 = ((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 4; c4 < ncol; c4 += nw * 4) {
    long cc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      long c = c0 + c4 + u;
      cc[u] = c < s0 ? c : c + skip;
    }
    for (int pl = 0; pl < nplans; pl++) {
      const int *plan = plans + (long)pl * PLAN_LEN;
      const int cnt = plan[PW * 4];
      const long dst = lane < cnt ? plan[lane] : 0, src = lane < cnt ? plan[PW * 2 + lane] : 0;
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (c4 + u < ncol && lane < cnt) v[u] = A[src + cc[u] * lda];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (c4 + u < ncol && lane < cnt) A[dst + cc[u] * lda] = v[u];
      // the next plan's loads see these stores
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
  }
}

template <typename T>
static void laswp_plans(T *A, long lda, long c0, long c1, long s0, long s1, const int *plans, int nplans,
                        hipStream_t st, int nbat = 1, BStride bs = BStride())
{
  long ncol = (c1 - c0) - (s1 - s0);
  if (ncol <= 0 || nplans <= 0) return;
  long nwaves = (ncol + 3) / 4;
  unsigned blocks = (unsigned)std::min<long>(4096, (nwaves + 3) / 4);
  hipLaunchKernelGGL(k_laswp_plans<T>, dim3(blocks, (unsigned)nbat), dim3(256), 0, st, A, lda, c0, c1, s0, s1, plans,
                     nplans, bs);
}

// row interchanges rows [k1,k2) (absolute, ipiv 1-based) on columns [c0,c1)
template <typename T>
__global__ void k_laswp(T *A, long lda, long c0, long c1, long k1, long k2, const int *ipiv)
{
  long c = c0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  for (long s = k1; s < k2; s++) {
    long p = ipiv[s] - 1;
    if (p != s) { T t = A[s + c * lda]; A[s + c * lda] = A[p + c * lda]; A[p + c * lda] = t; }
  }
}

template <typename T>
static void laswp(T *A, long lda, long c0, long c1, long k1, long k2, const int *ipiv, hipStream_t s)
{
  if (c1 <= c0 || k2 <= k1) return;
  hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((c1 - c0 + 255) / 256)), dim3(256), 0, s, A, lda, c0, c1, k1, k2, ipiv);
}

// B[w x ncol] <- L^{-1} B for the unit-lower w x w block of the panel
// (w <= PW): L and a 128-column tile of B staged in LDS with coalesced
// loads; each thread then carries one column in registers
template <typename T> struct TsmCols;
template <> struct TsmCols<double2> { static const int v = 64; };
template <> struct TsmCols<float2> { static const int v = 128; };
template <> struct TsmCols<double> { static const int v = 128; };

template <typename T>
__global__ void __launch_bounds__(TsmCols<T>::v)
k_trsm_small(const T *A, long lda, long r0, int w, T *B, long ldb, long ncol)
{
  constexpr int TSC = TsmCols<T>::v;
  __shared__ T L[PW][PW + 1];
  __shared__ T Bt[PW][TSC + 1];
  const int t = threadIdx.x;
  const long c0 = (long)blockIdx.x * TSC;
  {
    // L and the B tile staged through registers in batches of 8 loads
    constexpr int NL = PW * PW / TSC;
    const bool full = (w == PW && c0 + TSC <= ncol);
#pragma unroll
    for (int u0 = 0; u0 < NL; u0 += 8) {
      T lv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        int e = t + (u0 + u) * TSC, r = e % PW, q = e / PW;
        if (full || (r < w && q < w)) lv[u] = A[(r0 + r) + (r0 + q) * lda];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        int e = t + (u0 + u) * TSC, r = e % PW, q = e / PW;
        if (full || (r < w && q < w)) L[r][q] = lv[u];
      }
    }
#pragma unroll
    for (int u0 = 0; u0 < PW; u0 += 8) {
      T bv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        int e = t + (u0 + u) * TSC, r = e % PW, cc = e / PW;
        if (full || (r < w && c0 + cc < ncol)) bv[u] = B[r + (c0 + cc) * ldb];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        int e = t + (u0 + u) * TSC, r = e % PW, cc = e / PW;
        if (full || (r < w && c0 + cc < ncol)) Bt[r][cc] = bv[u];
      }
    }
  }
  __syncthreads();
  T x[PW];
#pragma unroll
  for (int r = 0; r < PW; r++) if (r < w) x[r] = Bt[r][t];
#pragma unroll
  for (int q = 0; q < PW; q++) {
    if (q < w) {
#pragma unroll
      for (int r = q + 1; r < PW; r++)
        if (r < w) x[r] = csub(x[r], cmul(L[r][q], x[q]));
    }
  }
#pragma unroll
  for (int r = 0; r < PW; r++) if (r < w) Bt[r][t] = x[r];
  __syncthreads();
  for (int e = t; e < PW * TSC; e += TSC) {
    int r = e % PW, cc = e / PW;
    if (r < w && c0 + cc < ncol) B[r + (c0 + cc) * ldb] = Bt[r][cc];
  }
}

template <typename T>
static void trsm_small(const T *A, long lda, long r0, int w, T *B, long ldb, long ncol, hipStream_t s)
{
  if (ncol <= 0 || w <= 0) return;
  int tok = prof_begin(s, std::is_same<T, double2>::value ? PK_TRSM_C128 : PK_TRSM_C64);
  constexpr int TSC = TsmCols<T>::v;
  hipLaunchKernelGGL(k_trsm_small<T>, dim3((unsigned)((ncol + TSC - 1) / TSC)), dim3(TSC), 0, s, A, lda, r0, w, B, ldb, ncol);
  prof_end(s, tok, 4.0 * (double)w * (double)w * (double)ncol);
}

// B[nb x ncol] <- L^{-1} B with the nb x nb unit-lower block at A(r0,r0),
// by PW-row steps: substitution on the diagonal step (trsm_small) and a
// GEMM update of the rows below it, as a blocked LAPACK trsm does
template <typename T>
static void trsm_block(const T *A, long lda, long r0, int nb, T *B, long ldb, long ncol, hipStream_t s)
{
  for (int i0 = 0; i0 < nb; i0 += PW) {
    int w = std::min(PW, nb - i0);
    trsm_small(A, lda, r0 + i0, w, B + i0, ldb, ncol, s);
    if (i0 + w < nb)
      gemm_sub((long)(nb - i0 - w), ncol, (long)w, A + (r0 + i0 + w) + (r0 + i0) * lda, lda, B + i0, ldb,
               B + i0 + w, ldb, s);
  }
}

// B[w x ncol] <- Linv B for a w x w (w <= PW) inverse of a unit-lower
// panel block, in place: each workgroup stages Linv and a 64-column tile
// of B in LDS; thread (row, column group) forms one row of 8 columns
template <typename T>
__global__ void __launch_bounds__(256)
k_trmm_strip(const T *Linv, int w, T *B, long ldb, long ncol)
{
  __shared__ T Ls[PW][PW + 1];
  __shared__ T Bt[PW][64 + 1];
  const int t = threadIdx.x;
  const long c0 = (long)blockIdx.x * 64;
  const bool full = (w == PW && c0 + 64 <= ncol);
  {
    T lv[4], bv[8];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int e = t + 256 * u, r = e & (PW - 1), k = e / PW;
      lv[u] = Linv[r + k * PW];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int e = t + 256 * u, r = e & (PW - 1), c = e / PW;
      if (full || (r < w && c0 + c < ncol)) bv[u] = B[r + (c0 + c) * ldb];
      else { bv[u] = zero_v<T>(); }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int e = t + 256 * u, r = e & (PW - 1), k = e / PW;
      Ls[r][k] = lv[u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int e = t + 256 * u, r = e & (PW - 1), c = e / PW;
      Bt[r][c] = bv[u];
    }
  }
  __syncthreads();
  const int r = t & (PW - 1), cg = t >> 5;
  T acc[8];
#pragma unroll
  for (int u = 0; u < 8; u++) { acc[u] = zero_v<T>(); }
#pragma unroll 8
  for (int k = 0; k < PW; k++) {
    const T a = Ls[r][k];
#pragma unroll
    for (int u = 0; u < 8; u++) acc[u] = cadd(acc[u], cmul(a, Bt[k][cg * 8 + u]));
  }
  if (r < w) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      long c = c0 + cg * 8 + u;
      if (c < ncol) B[r + c * ldb] = acc[u];
    }
  }
}

template <typename T>
static void trmm_strip(const T *Linv, int w, T *B, long ldb, long ncol, hipStream_t s)
{
  if (ncol <= 0 || w <= 0) return;
  int tok = prof_begin(s, std::is_same<T, double2>::value ? PK_TRSM_C128 : PK_TRSM_C64);
  hipLaunchKernelGGL(k_trmm_strip<T>, dim3((unsigned)((ncol + 63) / 64)), dim3(256), 0, s, Linv, w, B, ldb, ncol);
  prof_end(s, tok, 4.0 * (double)w * (double)w * (double)ncol);
}

// One panel's interchanges on the columns [c0, c1) outside the panel
// [s0, s1) and, on the columns right of the panel, the product of the
// panel block's unit-lower inverse with its w rows from r0 (what
// laswp_plans and trmm_strip do, in one launch): one wave per column.
// Lane e holds plan entry e (dst <- src, all sources read before any
// write); entries 0..w-1 are the rows r0 .. r0 + w - 1, so after the
// interchange lanes 0..w-1 hold the panel rows of the column.
__device__ inline float readlane_f(float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); }
__device__ inline double readlane_f(double v, int k)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFF), k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_swap_trmm(T *A, long lda, long c0, long c1, long s0, long s1, const int *plan, const T *Linv, int w, long r0,
            BStride bs)
{
  {
    const long bi = blockIdx.y;   // instance of a batch
    A += bi * bs.a;
    plan = bshift(plan, bi * bs.s);
    Linv = bshift(Linv, bi * bs.s);
  }
  const int lane = threadIdx.x & 63;
  const long ci = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long nleft = s0 - c0;
  if (ci >= (c1 - c0) - (s1 - s0)) return;
  const long c = ci < nleft ? c0 + ci : s1 + (ci - nleft);
  const int cnt = plan[PW * 4];
  T v = zero_v<T>();
  long dst = -1;
  if (lane < cnt) {
    dst = plan[lane];
    v = A[plan[PW * 2 + lane] + c * lda];
  }
  if (c >= s1) {
    // this lane's row of Linv, all loads in flight before the products
    T lr[PW];
    const int q = lane < w ? lane : 0;
#pragma unroll
    for (int k = 0; k < PW; k++) lr[k] = Linv[q + k * PW];
    T acc = zero_v<T>();
#pragma unroll
    for (int k = 0; k < PW; k++) {
      if (k < w) {
        T vk;
        vk.x = readlane_f(v.x, k);
        vk.y = readlane_f(v.y, k);
        acc = cadd(acc, cmul(lr[k], vk));
      }
    }
    if (lane < w) v = acc;
  }
  if (lane < cnt) A[dst + c * lda] = v;
}

template <typename T>
static void swap_trmm(T *A, long lda, long c0, long c1, long s0, long s1, const int *plan, const T *Linv, int w,
                      long r0, hipStream_t s, int nbat = 1, BStride bs = BStride())
{
  const long ncol = (c1 - c0) - (s1 - s0);
  if (ncol <= 0) return;
  int tok = prof_begin(s, std::is_same<T, double2>::value ? PK_TRSM_C128 : PK_TRSM_C64);
  hipLaunchKernelGGL(k_swap_trmm<T>, dim3((unsigned)((ncol + 3) / 4), (unsigned)nbat), dim3(256), 0, s, A, lda, c0, c1,
                     s0, s1, plan, Linv, w, r0, bs);
  prof_end(s, tok, 4.0 * (double)w * (double)w * (double)(c1 - s1));
}

// trsm_block with the panels' block inverses: per 32-row step a product
// with the step's inverse instead of substitution, then the GEMM update
template <typename T>
static void trsm_block_inv(const T *A, long lda, long r0, int nb, const T *linv, T *B, long ldb, long ncol,
                           hipStream_t s)
{
  for (int i0 = 0; i0 < nb; i0 += PW) {
    int w = std::min(PW, nb - i0);
    trmm_strip(linv + (long)(i0 / PW) * PW * PW, w, B + i0, ldb, ncol, s);
    if (i0 + w < nb)
      gemm_sub((long)(nb - i0 - w), ncol, (long)w, A + (r0 + i0 + w) + (r0 + i0) * lda, lda, B + i0, ldb,
               B + i0 + w, ldb, s);
  }
}

// B[jb x ncol] <- L^-1 B for the unit-lower jb x jb diagonal block of a
// 256-column block (jb <= 256) in one launch, with the panels' inverses:
// per 32-row panel i, B_i <- Linv_i B_i, then the rows below minus
// L(below, panel i) B_i (what trsm_block_inv does in 15 launches for the
// next block's columns).  One workgroup per column: wave v holds rows
// 64 v + lane (jb / 64 waves); the wave that holds panel i's rows forms x_i = Linv_i B_i
// by lane reads and passes it to the others through LDS, so the rows
// below are updated by four waves at once.
template <typename T>
__global__ void __launch_bounds__(256)
k_trsm_blk(const T *A, long lda, long r0, int jb, const T *linv, T *B, long ldb, long ncol, BStride bs)
{
  {
    const long bi = blockIdx.y;   // instance of a batch
    A += bi * bs.a;
    B += bi * bs.a;
    linv = bshift(linv, bi * bs.s);
  }
  __shared__ T xs[2][PW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long c = blockIdx.x;
  if (c >= ncol) return;
  const int row = lane + 64 * wv;
  T b = zero_v<T>();
  if (row < jb) b = B[row + c * ldb];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int i0 = PW * i;
    if (i0 < jb) {
      const int w = min(PW, jb - i0), si = i / 2, l0 = PW * (i & 1);
      if (wv == si) {
        // x = Linv_i b_i on the lanes holding the panel's rows
        const T *Li = linv + (long)i * PW * PW;
        const bool mine = lane >= l0 && lane < l0 + w;
        // this lane's row of Linv_i, all loads in flight before the products
        T lr[PW];
        const int q = mine ? lane - l0 : 0;
#pragma unroll
        for (int k = 0; k < PW; k++) lr[k] = Li[q + k * PW];
        T acc = zero_v<T>();
#pragma unroll
        for (int k = 0; k < PW; k++) {
          if (k < w) {
            T bk;
            bk.x = readlane_f(b.x, l0 + k);
            bk.y = readlane_f(b.y, l0 + k);
            acc = cadd(acc, cmul(lr[k], bk));
          }
        }
        if (mine) { b = acc; xs[i & 1][lane - l0] = acc; }
      }
      __syncthreads();
      // rows below the panel: minus L(rows, panel) x
      if (i0 + w < jb && row >= i0 + w && row < jb) {
        const T *Lr = A + (r0 + row) + (r0 + i0) * lda;
        T lv[PW];
#pragma unroll
        for (int k = 0; k < PW; k++) lv[k] = Lr[(k < w ? k : 0) * lda];
        T a = b;
#pragma unroll
        for (int k = 0; k < PW; k++)
          if (k < w) a = csub(a, cmul(lv[k], xs[i & 1][k]));
        b = a;
      }
    }
  }
  if (row < jb) B[row + c * ldb] = b;
}

template <typename T>
static void trsm_blk(const T *A, long lda, long r0, int jb, const T *linv, T *B, long ldb, long ncol, hipStream_t s,
                     int nbat = 1, BStride bs = BStride())
{
  if (ncol <= 0 || jb <= 0) return;
  int tok = prof_begin(s, std::is_same<T, double2>::value ? PK_TRSM_C128 : PK_TRSM_C64);
  // one wave per 64 rows of the block: two for a 128-column block
  const unsigned nthr = jb <= 128 ? 128u : 256u;
  hipLaunchKernelGGL(k_trsm_blk<T>, dim3((unsigned)ncol, (unsigned)nbat), dim3(nthr), 0, s, A, lda, r0, jb, linv, B,
                     ldb, ncol, bs);
  prof_end(s, tok, 4.0 * (double)jb * (double)jb * (double)ncol);
}

// second stream of the lookahead (default priority, non-blocking), one per
// device, created on first use
// second stream of the lookahead (default priority, non-blocking), one per
// device, created on first use.  (Keeping its kernels off a share of the
// CUs with a CU-masked stream was measured 1.7x slower on gfx950.)
static hipStream_t side_stream()
{
  static hipStream_t st[64] = {nullptr};
  int d = 0;
  hipGetDevice(&d);
  if (!st[d & 63]) hipStreamCreateWithFlags(&st[d & 63], hipStreamNonBlocking);
  return st[d & 63];
}

// Right-looking blocked LU with one block of lookahead.  Block b (nb
// columns from K) is factored in 32-column panels on the caller's stream
// (panel, its row interchanges inside the block, small TRSM and GEMM for
// the block's remaining columns); the next block's columns are then
// updated on the same stream, while the rest of the trailing matrix is
// updated on the side stream.  The next block's panels therefore run
// while the big GEMM of this block does.  Row interchanges on columns
// left of a block are applied at the end (nothing reads those rows in
// between), so the side stream's GEMM never sees its L21 rows move.
template <typename T>
int lu_factor(T *A, long n, long lda, int *d_ipiv, int *info, hipStream_t stream, int nb, int fast)
{
  if (n < 0 || lda < std::max(1L, n)) return -1;
  if (n == 0) return 0;
  if (nb > 256) nb = 256;
  if (nb % PW) nb = 256;
  const int ib = PW;
  constexpr int NR = PanelRows<T>::n;
  const long Gmax = (n + NR - 1) / NR;
  // the mixed solver's complex64 factorization pivots by tournament
  // (SCUFFX_MIXED_PIVOTING=partial, read at every factorization: row by
  // row, as the complex128 path)
  const char *piv_env = getenv("SCUFFX_MIXED_PIVOTING");
  const bool partial_env = piv_env && !strcmp(piv_env, "partial");
  if (IsReal<T>::v) fast = 0;     // the explicit-inverse block rows are a complex64 option
  const bool tourn = fast && std::is_same<T, float2>::value && !partial_env;
  const bool g3m = tourn && mixed_gemm_3m();
  const void *panel_fn = tourn ? (const void *)k_panel_ts : (const void *)k_panel<T>;
  int dev = 0, ncu = 0, perCU = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, panel_fn, NR, 0);
  // the tournament launches one more workgroup (it inverts the diagonal block)
  if ((long)std::max(1, ncu) * std::max(1, perCU) < Gmax + (tourn ? 1 : 0)) return -5;   // not co-resident
  if (Gmax > 64 * KEYS_PER_LANE) return -5;                           // more workgroups than polled keys
  const long npanels = (n + PW - 1) / PW;
  // panel scratch, carved from one workspace: barrier words (3 cache
  // lines), info, all panel plans, candidate values / indices / rows,
  // diagonal rows
  PanelScratch S;
  size_t off_info = 96 * sizeof(unsigned), off_plan = off_info + 64;
  size_t off_cv = off_plan + 256 * ((sizeof(int) * PLAN_LEN * npanels + 255) / 256);
  size_t off_ci = off_cv + 256 * ((sizeof(double) * 2 * Gmax + 255) / 256);
  size_t off_cr = off_ci + 256 * ((sizeof(unsigned long long) * 2 * Gmax + 255) / 256);
  size_t off_rc = off_cr + 256 * ((sizeof(T) * 2 * Gmax * PW + 255) / 256);
  size_t off_li = off_rc + 256 * ((sizeof(T) * 2 * PW + 255) / 256);
  size_t off_ts = off_li + 256 * (((fast ? sizeof(T) * PW * PW * npanels : 0) + 255) / 256);
  const size_t ts_words = tourn ? (size_t)ts_root_off(ts_levels(Gmax), Gmax) + TS_ROOT : 0;
  size_t total = off_ts + sizeof(unsigned long long) * ts_words;
  char *wsb = (char *)ws_get(WS_PANEL, total);
  if (!wsb) return -2;
  S.linv = fast ? (void *)(wsb + off_li) : nullptr;
  S.ts = tourn ? (unsigned long long *)(wsb + off_ts) : nullptr;
  S.bar = (unsigned *)wsb;
  int *dinfo = (int *)(wsb + off_info);
  int *plans = (int *)(wsb + off_plan);
  S.candv = (double *)(wsb + off_cv);
  S.ckey = (unsigned long long *)(wsb + off_ci);
  S.candrow = (void *)(wsb + off_cr);
  S.rowc = (void *)(wsb + off_rc);
  // The trailing update runs beside the next panels on a second stream,
  // unless the panel needs more than half of the resident workgroup slots
  // of its kernel: then the GEMM's workgroups, dispatched without pause
  // on the other stream, could keep part of the panel's grid waiting until
  // its bounded waits expire (seen at n = 61440 complex128: 480 panel
  // workgroups of 2 per CU).  Everything then runs on the caller's stream.
  const bool overlap = 2 * (Gmax + (tourn ? 1 : 0)) <= (long)std::max(1, ncu) * std::max(1, perCU);
  hipStream_t side = overlap ? side_stream() : stream;
  hipEvent_t evP = nullptr, evR = nullptr;
  hipEventCreateWithFlags(&evP, hipEventDisableTiming);
  hipEventCreateWithFlags(&evR, hipEventDisableTiming);
  hipMemsetAsync(dinfo, 0, sizeof(int), stream);
  hipMemsetAsync(S.bar, 0, sizeof(unsigned) * 96, stream);
  // candidate keys start at sequence 0, which no column uses
  hipMemsetAsync(S.ckey, 0, sizeof(unsigned long long) * 2 * Gmax, stream);
  // tournament sets and flags likewise
  if (tourn) hipMemsetAsync(S.ts, 0, sizeof(unsigned long long) * ts_words, stream);
  // the side stream starts after the caller's stream has the matrix ready
  hipEventRecord(evR, stream);
  hipStreamWaitEvent(side, evR, 0);
  int rc = 0;
  int gmaxi = (int)Gmax;
  for (long K = 0; K < n && rc == 0; K += nb) {
    const long jb = std::min<long>(nb, n - K);
    const long p0 = K / PW;                 // first panel of the block
    for (long j0 = K; j0 < K + jb; j0 += ib) {
      int w = (int)std::min<long>(ib, K + jb - j0);
      long m = n - j0;
      long G = (m + NR - 1) / NR;
      long r0 = j0;
      S.plan = plans + (j0 / PW) * PLAN_LEN;
      // column sequence numbers: fresh for every panel of this factorization
      unsigned seq0 = (unsigned)(j0 + 1);
      BStride one;
      void *args[] = {(void *)&A, (void *)&lda, (void *)&r0, (void *)&m, (void *)&w, (void *)&d_ipiv,
                      (void *)&dinfo, (void *)&S, (void *)&gmaxi, (void *)&seq0, (void *)&one};
      int tok = prof_begin(stream, IsReal<T>::v ? PK_PANEL_F64
                                   : std::is_same<T, double2>::value ? PK_PANEL_C128 : PK_PANEL_C64);
      // the tournament panel has no grid barrier, only bounded waits along
      // its merge tree: an ordinary launch keeps it on the stream's own
      // hardware queue (a cooperative launch goes through a queue of its
      // own, with a cross-queue handoff before and after every panel)
      hipError_t e = tourn ? hipLaunchKernel(panel_fn, dim3((unsigned)G + 1), dim3(NR), args, 0, stream)
                           : hipLaunchCooperativeKernel(panel_fn, dim3((unsigned)G), dim3(NR), args, 0, stream);
      // algorithmic flops of the m x w panel: per column, one complex
      // reciprocal-multiply per row below it and a complex rank-1 update
      double pflops = 0;
      for (int j = 0; j < w; j++) pflops += (double)(m - j - 1) * (6.0 + 8.0 * (double)(w - j - 1));
      prof_end(stream, tok, pflops);
      if (e != hipSuccess) { rc = -3; break; }
      // this panel's interchanges on the block's other columns (fast: with
      // the product by the panel block's inverse on the columns right of it)
      long nr = K + jb - (j0 + w);
      if (fast) {
        if constexpr (!IsReal<T>::v)
          swap_trmm(A, lda, K, K + jb, j0, j0 + w, S.plan, (const T *)S.linv + (j0 / PW) * PW * PW, w, j0, stream);
      } else
        laswp_plans(A, lda, K, K + jb, j0, j0 + w, S.plan, 1, stream);
      if (nr > 0) {
        if (!fast) trsm_small(A, lda, j0, w, A + j0 + (j0 + w) * lda, lda, nr, stream);
        gemm_sub(n - j0 - w, nr, (long)w, A + (j0 + w) + j0 * lda, lda, A + j0 + (j0 + w) * lda, lda,
                 A + (j0 + w) + (j0 + w) * lda, lda, stream, 0, 1, 0, g3m);
      }
    }
    if (rc) break;
    const int bplans = (int)((jb + PW - 1) / PW);
    const long c1 = K + jb;                               // first column right of the block
    const long nnext = std::min<long>(nb, n - c1), c2 = c1 + nnext;
    hipEventRecord(evP, stream);
    // the next block's columns (caller's stream), after the side stream
    // has finished the previous block's update of them
    if (nnext > 0) {
      hipStreamWaitEvent(stream, evR, 0);
      laswp_plans(A, lda, c1, c2, 0, 0, plans + p0 * PLAN_LEN, bplans, stream);
      if (fast) {
        if constexpr (!IsReal<T>::v)
          trsm_blk(A, lda, K, (int)jb, (const T *)S.linv + p0 * PW * PW, A + K + c1 * lda, lda, nnext, stream);
      } else
        trsm_block(A, lda, K, (int)jb, A + K + c1 * lda, lda, nnext, stream);
      gemm_sub(n - c1, nnext, jb, A + c1 + K * lda, lda, A + K + c1 * lda, lda, A + c1 + c1 * lda, lda, stream, 0, 1, 0,
               g3m);
    }
    // the rest of the trailing matrix (side stream)
    if (c2 < n) {
      hipStreamWaitEvent(side, evP, 0);
      laswp_plans(A, lda, c2, n, 0, 0, plans + p0 * PLAN_LEN, bplans, side);
      if (fast) {
        if constexpr (!IsReal<T>::v)
          trsm_block_inv(A, lda, K, (int)jb, (const T *)S.linv + p0 * PW * PW, A + K + c2 * lda, lda, n - c2, side);
      } else
        trsm_block(A, lda, K, (int)jb, A + K + c2 * lda, lda, n - c2, side);
      // beside tournament panels: at most three trailing-GEMM workgroups
      // per CU (44 KB of LDS each), so that a panel workgroup (168 VGPRs
      // per wave, 12 KB) finds room on every CU instead of waiting for
      // GEMM workgroups to retire
      const unsigned pad = tourn ? 44u * 1024u - (unsigned)(4 * GBK * (GBM + 4) * sizeof(float)) : 0u;
      gemm_sub(n - c1, n - c2, jb, A + c1 + K * lda, lda, A + K + c2 * lda, lda, A + c1 + c2 * lda, lda, side, pad, 1,
               0, g3m);
      hipEventRecord(evR, side);
    }
    // this block's interchanges on the columns left of it.  Nothing reads
    // those rows again during the factorization (the trailing updates read
    // only the current block's L21), so with two streams they run on the
    // side stream behind this block's update, in block order, instead of
    // all at the end of the caller's stream
    if (side != stream && K > 0) {
      if (c2 >= n) hipStreamWaitEvent(side, evP, 0);   // no trailing update queued for this block
      laswp_plans(A, lda, 0, K, 0, 0, plans + p0 * PLAN_LEN, bplans, side);
    }
  }
  // join the side stream (one stream: the deferred interchanges on the
  // columns left of each block, now)
  hipEventRecord(evR, side);
  hipStreamWaitEvent(stream, evR, 0);
  if (rc == 0 && side == stream)
    for (long K = nb; K < n; K += nb) {
      const long jb = std::min<long>(nb, n - K);
      laswp_plans(A, lda, 0, K, 0, 0, plans + (K / PW) * PLAN_LEN, (int)((jb + PW - 1) / PW), stream);
    }
  unsigned tmo = 0;
  hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, stream);
  hipMemcpyAsync(&tmo, S.bar + 64, sizeof(unsigned), hipMemcpyDeviceToHost, stream);
  hipStreamSynchronize(stream);
  hipStreamSynchronize(side);
  hipEventDestroy(evP);
  hipEventDestroy(evR);
  if (prof_enabled()) prof_collect();
  if (tmo && rc == 0) rc = tmo == 1 ? -6 : -3;
  if (rc == 0 && hipGetLastError() != hipSuccess) rc = -3;
  return rc;
}

// ---------------------------------------------------------------------
// triangular solves.  Blocked for trans='N': a diagonal block is solved
// by one workgroup per right-hand side, the rest updated by the GEMM.
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_trsv_block(const T *A, long lda, long r0, int bs, T *B, long ldb, int lower, int unit)
{
  __shared__ T x[256];
  T *b = B + (long)blockIdx.x * ldb + r0;
  const int tid = threadIdx.x;
  if (tid < bs) x[tid] = b[tid];
  __syncthreads();
  if (lower) {
    for (int i = 0; i < bs; i++) {
      if (!unit && tid == 0) x[i] = cdiv(x[i], A[(r0 + i) + (r0 + i) * lda]);
      __syncthreads();
      T xi = x[i];
      if (tid > i && tid < bs) x[tid] = csub(x[tid], cmul(A[(r0 + tid) + (r0 + i) * lda], xi));
      __syncthreads();
    }
  } else {
    for (int i = bs - 1; i >= 0; i--) {
      if (!unit && tid == 0) x[i] = cdiv(x[i], A[(r0 + i) + (r0 + i) * lda]);
      __syncthreads();
      T xi = x[i];
      if (tid < i) x[tid] = csub(x[tid], cmul(A[(r0 + tid) + (r0 + i) * lda], xi));
      __syncthreads();
    }
  }
  if (tid < bs) b[tid] = x[tid];
}

// transposed / conjugate-transposed solves, one workgroup per column
template <typename T>
__global__ void __launch_bounds__(256)
k_trsv_trans(const T *A, long lda, long n, T *B, long ldb, int upper_first, int conj_)
{
  T *b = B + (long)blockIdx.x * ldb;
  __shared__ T red[256];
  const int tid = threadIdx.x;
  // solve U^T y = b (forward), then L^T x = y (backward)
  for (int pass = 0; pass < 2; pass++) {
    bool up = (pass == 0);
    for (long s = 0; s < n; s++) {
      long i = up ? s : n - 1 - s;
      T acc = zero_v<T>();
      if (up) { for (long r = tid; r < i; r += 256) { T a = conj_if(A[r + i * lda], conj_); mac(acc, a, b[r]); } }
      else { for (long r = i + 1 + tid; r < n; r += 256) { T a = conj_if(A[r + i * lda], conj_); mac(acc, a, b[r]); } }
      red[tid] = acc;
      __syncthreads();
      for (int q = 128; q > 0; q >>= 1) { if (tid < q) red[tid] = cadd(red[tid], red[tid + q]); __syncthreads(); }
      if (tid == 0) {
        T v = csub(b[i], red[0]);
        if (up) { T d = conj_if(A[i + i * lda], conj_); v = cdiv(v, d); }
        b[i] = v;
      }
      __syncthreads();
    }
  }
  (void)upper_first;
}

template <typename T>
__global__ void k_laswp_rev(T *B, long ldb, long n, long nrhs, const int *ipiv)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  for (long s = n - 1; s >= 0; s--) {
    long p = ipiv[s] - 1;
    if (p != s) { T t = B[s + c * ldb]; B[s + c * ldb] = B[p + c * ldb]; B[p + c * ldb] = t; }
  }
}

// The row interchanges of ipiv (LAPACK order, each row s swapped with a
// row p >= s) composed into one permutation: perm[s] is the row whose
// content ends up in row s.  One wave walks the interchanges in order on
// an LDS copy of the permutation; a ballot over 64 pivots at a time skips
// the rows that stay (most of them for a BEM matrix).  The right-hand
// sides are then gathered (forward) or scattered (reverse) in parallel,
// instead of one thread per column walking all n interchanges in memory.
#define PERM_LDS_MAX 20480
__global__ void __launch_bounds__(64) k_perm_build(const int *ipiv, long n, int *perm, long sp, long sperm)
{
  ipiv += (long)blockIdx.x * sp;   // instance blockIdx.x of a batch
  perm += (long)blockIdx.x * sperm;
  __shared__ int P[PERM_LDS_MAX], V[PERM_LDS_MAX];
  const int lane = threadIdx.x;
  for (long i = lane; i < n; i += 64) { P[i] = (int)i; V[i] = ipiv[i] - 1; }
  __syncthreads();
  for (long s0 = 0; s0 < n; s0 += 64) {
    const long s = s0 + lane;
    const int p = s < n ? V[s] : (int)s;
    unsigned long long m = __ballot(s < n && p != s && p > s && p < n);   // interchanges stay below row n
    while (m) {
      const int b = __ffsll(m) - 1;
      m &= m - 1;
      const int pp = __shfl(p, b);
      // one lane swaps; the wave's LDS operations complete in order
      if (lane == 0) { const int a = P[s0 + b]; P[s0 + b] = P[pp]; P[pp] = a; }
    }
  }
  __syncthreads();
  for (long i = lane; i < n; i += 64) perm[i] = P[i];
}

// forward: out[s] = in[perm[s]]; reverse: out[perm[s]] = in[s]
template <typename T, bool FWD>
__global__ void __launch_bounds__(256) k_perm_apply(const T *in, long ldi, T *out, long ldo, const int *perm, long n,
                                                    long sin, long sout, long sperm)
{
  {
    const long z = blockIdx.z;   // instance of a batch
    in += z * sin; out += z * sout; perm += z * sperm;
  }
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const long c = blockIdx.y;
  const long p = perm[s];
  if (FWD) out[s + c * ldo] = in[p + c * ldi];
  else out[p + c * ldo] = in[s + c * ldi];
}

// applies the interchanges of ipiv to the nrhs columns of B (forward, as
// laswp with increment 1) or their inverse (reverse order); tmp holds n x
// nrhs elements; perm_ready: perm_buf already holds the composed
// permutation of ipiv
template <typename T>
static void apply_row_perm(T *B, long ldb, long n, long nrhs, const int *ipiv, bool fwd, T *tmp, hipStream_t s,
                           int *perm_buf = nullptr, bool perm_ready = false);

// Triangular solves with few right-hand sides, in 64-row block steps.
// The diagonal 64x64 blocks of L (unit lower) and U (upper) are inverted
// first, all blocks at once (k_diag_inv), so every block step is a small
// matrix-vector product instead of a 64-step serial substitution.
#define TB 64
template <typename T, bool LOWER>
__global__ void __launch_bounds__(TB)
k_diag_inv(const T *LU, long lda, long n, T *inv, long sa, long sinv)
{
  {
    const long z = blockIdx.z;   // instance of a batch
    LU += z * sa; inv += z * sinv;
  }
  __shared__ T D[TB][TB + 1];
  __shared__ T X[TB][TB + 1];       // X[r][j]: column j of the inverse
  const int t = threadIdx.x;
  const long kb = (long)blockIdx.x * TB;
  const int bs = (int)std::min<long>(TB, n - kb);
  for (int q = 0; q < TB; q++) {
    T v;
    if (t < bs && q < bs) v = LU[(kb + t) + (kb + q) * lda];
    else v = (t == q) ? unit_v<T>() : zero_v<T>();
    D[t][q] = v;
  }
  __syncthreads();
  const int j = t;
  if (LOWER) {
    for (int r = 0; r < j; r++) { X[r][j] = zero_v<T>(); }
    X[j][j] = unit_v<T>();
    for (int r = j + 1; r < TB; r++) {
      T acc = zero_v<T>();
      for (int k = j; k < r; k++) acc = csub(acc, cmul(D[r][k], X[k][j]));
      X[r][j] = acc;
    }
  } else {
    for (int r = j + 1; r < TB; r++) { X[r][j] = zero_v<T>(); }
    X[j][j] = crecip(D[j][j]);
    for (int r = j - 1; r >= 0; r--) {
      T acc = zero_v<T>();
      for (int k = r + 1; k <= j; k++) acc = csub(acc, cmul(D[r][k], X[k][j]));
      X[r][j] = cdiv(acc, D[r][r]);
    }
  }
  __syncthreads();
  T *o = inv + (long)blockIdx.x * TB * TB;
  for (int q = 0; q < TB; q++) o[t + q * TB] = X[t][q];
}

// one block step: x_b = inv(D_b) r_b (every workgroup, from L2), the
// solution written to Xout by workgroup 0, and the rows outside the block
// updated with x_b (LOWER: rows below, else rows above); blockIdx.y = rhs
template <typename T, bool LOWER>
__global__ void __launch_bounds__(256)
k_trsv_inv(const T *LU, long lda, long n, long kb, int bs, const T *invb, T *Bin, long ldb, T *Xout, long ldx,
           long sa, long sinv, long sb, long sx)
{
  {
    const long z = blockIdx.z;   // instance of a batch
    LU += z * sa; invb += z * sinv; Bin += z * sb; Xout += z * sx;
  }
  __shared__ T rs[TB], xs[TB];
  const int t = threadIdx.x;
  T *b = Bin + (long)blockIdx.y * ldb;
  T *xo = Xout + (long)blockIdx.y * ldx;
  if (t < TB) {
    if (t < bs) rs[t] = b[kb + t]; else { rs[t] = zero_v<T>(); }
  }
  __syncthreads();
  if (t < TB) {
    T acc = zero_v<T>();
#pragma unroll
    for (int k0 = 0; k0 < TB; k0 += 16) {
      T a[16];
#pragma unroll
      for (int u = 0; u < 16; u++) a[u] = invb[t + (k0 + u) * TB];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        T v = rs[k0 + u];
        mac(acc, a[u], v);
      }
    }
    xs[t] = acc;
  }
  __syncthreads();
  if (blockIdx.x == 0 && t < bs) xo[kb + t] = xs[t];
  const long row = LOWER ? kb + bs + (long)blockIdx.x * 256 + t : (long)blockIdx.x * 256 + t;
  const long rend = LOWER ? n : kb;
  if (row < rend) {
    T acc = zero_v<T>();
    const T *Lr = LU + row + kb * lda;
    T bo = b[row];
    if (bs == TB) {
      // full block: 16 independent loads in flight per batch
#pragma unroll
      for (int i0 = 0; i0 < TB; i0 += 16) {
        T a[16];
#pragma unroll
        for (int u = 0; u < 16; u++) a[u] = Lr[(long)(i0 + u) * lda];
#pragma unroll
        for (int u = 0; u < 16; u++) {
          T v = xs[i0 + u];
          mac(acc, a[u], v);
        }
      }
    } else {
      for (int i = 0; i < bs; i++) {
        T a = Lr[(long)i * lda], v = xs[i];
        mac(acc, a, v);
      }
    }
    b[row] = csub(bo, acc);
  }
}

// One right-hand side, all block steps in one cooperative launch.
// Workgroup g owns the 64 rows of block g: it subtracts the off-diagonal
// blocks' products with every earlier solution block as that block is
// published (forward: blocks k < g of L; backward: k > g of U), then
// applies the inverse diagonal block and publishes its own 64 values.  The
// output starts filled with a sentinel (a NaN bit pattern in every 64-bit
// word that no solve produces), and every entry is written once, by an
// agent-scope relaxed atomic store: a reader that sees a word other than
// the sentinel sees its final value, so no flag or store drain is needed
// (one memory round trip per block step).  Every wait is bounded (error
// word, as in the panel).  Thread t handles row t % 64 and the 16 columns
// 16 (t / 64) .. of a block; the four partial sums meet in LDS.  The next
// block's 16 values per thread are loaded before the wait.  The forward
// solve, which reads its right-hand side b once per row, leaves the
// sentinel in b (mark), where the backward solve writes its output.
#define SCUFFX_CHAIN_SENT 0x7FF4DEADBEEF0003ULL

// a loaded value is final once none of its 64-bit words is the sentinel
template <typename T>
__device__ inline bool chain_val_set(const T &v)
{
  unsigned long long w[sizeof(T) / 8];
  __builtin_memcpy(w, &v, sizeof(T));
  bool set = true;
#pragma unroll
  for (int j = 0; j < (int)(sizeof(T) / 8); j++) set = set && (w[j] != SCUFFX_CHAIN_SENT);
  return set;
}

template <typename T>
__device__ inline void chain_mark(T *p)
{
  unsigned long long *w = (unsigned long long *)p;
#pragma unroll
  for (int j = 0; j < (int)(sizeof(T) / 8); j++) st_coh(w + j, SCUFFX_CHAIN_SENT);
}

template <typename T>
__global__ void k_chain_fill(T *x, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    unsigned long long *w = (unsigned long long *)(x + i);
#pragma unroll
    for (int j = 0; j < (int)(sizeof(T) / 8); j++) w[j] = SCUFFX_CHAIN_SENT;
  }
}

template <typename T, bool LOWER>
__global__ void __launch_bounds__(256)
k_trsv_chain(const T *LU, long lda, long n, const T *inv, T *bin, T *xout, int mark, unsigned *err)
{
  __shared__ T part[4][TB];
  __shared__ T bx[TB];
  const int t = threadIdx.x, r = t & (TB - 1), q = t >> 6;
  const long nblk = (n + TB - 1) / TB;
  const long g = blockIdx.x;
  const long row = g * TB + r;
  const int bs = (int)std::min<long>(TB, n - g * TB);
  T b = zero_v<T>();
  if (q == 0 && r < bs) {
    b = bin[row];
    if (mark) chain_mark(bin + row);
  }
  // blocks k in solve order before g
  const long nk = LOWER ? g : nblk - 1 - g;
  auto blk = [&](long i) { return LOWER ? i : nblk - 1 - i; };
  auto load = [&](long k, T (&a)[16]) {
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const long col = k * TB + q * 16 + u;
      if (r < bs && col < n) a[u] = LU[row + col * lda];
      else { a[u] = zero_v<T>(); }
    }
  };
  // the inverse diagonal block, in the same registers once no block is left
  const T *ib = inv + g * TB * TB;
  auto load_inv = [&](T (&m)[16]) {
#pragma unroll
    for (int u = 0; u < 16; u++) m[u] = ib[r + (long)(q * 16 + u) * TB];
  };
  T a[16];
  if (nk > 0) load(blk(0), a);
  else load_inv(a);
  for (long i = 0; i < nk; i++) {
    const long k = blk(i);
    // this thread's 16 values of block k, once published (bounded)
    // (all 16 loads in flight at once; the whole set again while any
    // still holds the sentinel)
    T x[16];
    unsigned spins = 0;
    for (;;) {
      bool ready = true;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const long col = k * TB + q * 16 + u;
        x[u] = col < n ? ld_coh(xout + col) : zero_v<T>();
      }
#pragma unroll
      for (int u = 0; u < 16; u++) ready = ready && chain_val_set(x[u]);
      if (ready) break;
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (++spins > GBAR_SPIN) { __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
      __builtin_amdgcn_s_sleep(1);
    }
    T acc = zero_v<T>();
#pragma unroll
    for (int u = 0; u < 16; u++) {
      mac(acc, a[u], x[u]);
    }
    if (i + 1 < nk) load(blk(i + 1), a);
    else load_inv(a);
    part[q][r] = acc;
    __syncthreads();
    if (q == 0) {
      T s4 = part[0][r];
      s4 = cadd(s4, part[1][r]); s4 = cadd(s4, part[2][r]); s4 = cadd(s4, part[3][r]);
      b = csub(b, s4);
    }
    __syncthreads();
  }
  // own block: x_g = inv(D_g) b_g
  if (q == 0) bx[r] = b;
  __syncthreads();
  T acc = zero_v<T>();
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const T m = a[u], v = bx[q * 16 + u];
    mac(acc, m, v);
  }
  part[q][r] = acc;
  __syncthreads();
  if (q == 0 && r < bs) {
    T xg = part[0][r];
    xg = cadd(xg, part[1][r]); xg = cadd(xg, part[2][r]); xg = cadd(xg, part[3][r]);
    st_coh(xout + row, xg);
  }
}

__global__ void k_zero_u64(unsigned long long *w, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0ull;
}

// lower (forward) then upper (backward) solve of one right-hand side in
// two cooperative launches; returns -1 if it cannot run (grid not
// co-resident, workspace), -6 on a time-out
template <typename T>
static int trsv_chain(const T *LU, long lda, long n, const T *invL, const T *invU, T *B, T *W, hipStream_t s,
                      unsigned long long *words)
{
  const long nblk = (n + TB - 1) / TB;
  static int ncu = -1, perCU = 0;
  if (ncu < 0) {
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void *)k_trsv_chain<T, true>, 256, 0);
  }
  if ((long)ncu * perCU < nblk) return -1;
  char *w = words ? (char *)words : (char *)ws_get(WS_CHAIN, sizeof(unsigned long long) * (2 * nblk + 8));
  if (!w) return -1;
  unsigned *err = (unsigned *)w;
  // cleared and filled by kernels (not memset nodes, so that a captured
  // sequence does the same on every replay)
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, (unsigned long long *)w, 8L);
  hipLaunchKernelGGL(k_chain_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, W, n);
  int mark = 1, nomark = 0;
  void *al[] = {(void *)&LU, (void *)&lda, (void *)&n, (void *)&invL, (void *)&B, (void *)&W, (void *)&mark,
                (void *)&err};
  if (hipLaunchCooperativeKernel((const void *)k_trsv_chain<T, true>, dim3((unsigned)nblk), dim3(256), al, 0, s) !=
      hipSuccess)
    return -3;
  void *au[] = {(void *)&LU, (void *)&lda, (void *)&n, (void *)&invU, (void *)&W, (void *)&B, (void *)&nomark,
                (void *)&err};
  if (hipLaunchCooperativeKernel((const void *)k_trsv_chain<T, false>, dim3((unsigned)nblk), dim3(256), au, 0, s) !=
      hipSuccess)
    return -3;
  unsigned tmo = 0;
  hipMemcpyAsync(&tmo, err, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  return tmo ? -6 : 0;
}

template <typename T>
static void apply_row_perm(T *B, long ldb, long n, long nrhs, const int *ipiv, bool fwd, T *tmp, hipStream_t s,
                           int *perm_buf, bool perm_ready)
{
  int *perm = n > PERM_LDS_MAX ? nullptr : perm_buf ? perm_buf : (int *)ws_get(WS_PERM, sizeof(int) * n);
  if (!perm || !tmp) {
    // beyond the LDS copy: the interchanges one after another, per column
    if (fwd) laswp(B, ldb, 0, nrhs, 0, n, ipiv, s);
    else hipLaunchKernelGGL(k_laswp_rev<T>, dim3((unsigned)((nrhs + 63) / 64)), dim3(64), 0, s, B, ldb, n, nrhs, ipiv);
    return;
  }
  if (!(perm_ready && perm == perm_buf)) hipLaunchKernelGGL(k_perm_build, dim3(1), dim3(64), 0, s, ipiv, n, perm, 0L, 0L);
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)nrhs);
  if (fwd) hipLaunchKernelGGL((k_perm_apply<T, true>), grid, dim3(256), 0, s, B, ldb, tmp, n, perm, n, 0L, 0L, 0L);
  else hipLaunchKernelGGL((k_perm_apply<T, false>), grid, dim3(256), 0, s, B, ldb, tmp, n, perm, n, 0L, 0L, 0L);
  hipMemcpy2DAsync(B, sizeof(T) * ldb, tmp, sizeof(T) * n, sizeof(T) * n, nrhs, hipMemcpyDeviceToDevice, s);
}

template <typename T>
size_t solve_ws_bytes(long n)
{
  auto r256 = [](size_t b) { return 256 * ((b + 255) / 256); };
  const long nblk = (n + TB - 1) / TB;
  return r256(sizeof(T) * n) + 2 * r256(sizeof(T) * nblk * TB * TB) + r256(sizeof(int) * n) +
         r256(sizeof(unsigned long long) * (2 * nblk + 8));
}

template <typename T>
SolveWs<T> solve_ws_carve(char *base, long n)
{
  auto r256 = [](size_t b) { return 256 * ((b + 255) / 256); };
  const long nblk = (n + TB - 1) / TB;
  SolveWs<T> w;
  w.W = (T *)base; base += r256(sizeof(T) * n);
  w.invL = (T *)base; base += r256(sizeof(T) * nblk * TB * TB);
  w.invU = (T *)base; base += r256(sizeof(T) * nblk * TB * TB);
  w.perm = (int *)base; base += r256(sizeof(int) * n);
  w.chain = (unsigned long long *)base;
  return w;
}

template <typename T>
int lu_solve(const T *LU, long n, long lda, const int *d_ipiv, char trans, long nrhs, T *B, long ldb,
             hipStream_t stream, SolveWs<T> *ws)
{
  if (n < 0 || nrhs < 0 || lda < std::max(1L, n) || ldb < std::max(1L, n)) return -1;
  if (n == 0 || nrhs == 0) return 0;
  if (ws && !((trans == 'N' || trans == 'n') && nrhs == 1)) return -1;
  if ((trans == 'N' || trans == 'n') && nrhs <= 16) {
    const long nblk = (n + TB - 1) / TB;
    T *W = ws ? ws->W : (T *)ws_get(WS_SOLVE_W, sizeof(T) * n * nrhs);
    T *invL = ws ? ws->invL : (T *)ws_get(WS_SOLVE_INVL, sizeof(T) * nblk * TB * TB);
    T *invU = ws ? ws->invU : (T *)ws_get(WS_SOLVE_INVU, sizeof(T) * nblk * TB * TB);
    if (!W || !invL || !invU) return -2;
    // a solve after the first with the same factors reuses the block
    // inverses and the composed permutation
    const bool cached = ws && ws->inv_ready;
    if (!cached) {
      hipLaunchKernelGGL((k_diag_inv<T, true>), dim3((unsigned)nblk), dim3(TB), 0, stream, LU, lda, n, invL, 0L, 0L);
      hipLaunchKernelGGL((k_diag_inv<T, false>), dim3((unsigned)nblk), dim3(TB), 0, stream, LU, lda, n, invU, 0L, 0L);
      if (ws) ws->inv_ready = true;
    }
    apply_row_perm(B, ldb, n, nrhs, d_ipiv, true, W, stream, ws ? ws->perm : nullptr, cached);
    if (nrhs == 1) {
      int rc = trsv_chain(LU, lda, n, invL, invU, B, W, stream, ws ? ws->chain : nullptr);
      if (rc != -1) return rc;
    }
    // L y = P b: y accumulates in W, B holds the running right-hand side
    for (long kb = 0; kb < n; kb += TB) {
      int bs = (int)std::min<long>(TB, n - kb);
      long rows = n - kb - bs;
      dim3 grid((unsigned)std::max<long>(1, (rows + 255) / 256), (unsigned)nrhs);
      hipLaunchKernelGGL((k_trsv_inv<T, true>), grid, dim3(256), 0, stream, LU, lda, n, kb, bs,
                         invL + (kb / TB) * TB * TB, B, ldb, W, n, 0L, 0L, 0L, 0L);
    }
    // U x = y: W is updated in place, the solution lands in B
    long last = ((n - 1) / TB) * TB;
    for (long kb = last; kb >= 0; kb -= TB) {
      int bs = (int)std::min<long>(TB, n - kb);
      dim3 grid((unsigned)std::max<long>(1, (kb + 255) / 256), (unsigned)nrhs);
      hipLaunchKernelGGL((k_trsv_inv<T, false>), grid, dim3(256), 0, stream, LU, lda, n, kb, bs,
                         invU + (kb / TB) * TB * TB, W, n, B, ldb, 0L, 0L, 0L, 0L);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
  T *tmp = (T *)ws_get(WS_PERM_TMP, sizeof(T) * n * nrhs);
  if (trans == 'N' || trans == 'n') {
    apply_row_perm(B, ldb, n, nrhs, d_ipiv, true, tmp, stream);
    const int bs = 256;
    for (long kb = 0; kb < n; kb += bs) {
      int w = (int)std::min<long>(bs, n - kb);
      hipLaunchKernelGGL(k_trsv_block<T>, dim3((unsigned)nrhs), dim3(256), 0, stream, LU, lda, kb, w, B, ldb, 1, 1);
      if (kb + w < n) gemm_sub(n - kb - w, nrhs, (long)w, LU + (kb + w) + kb * lda, lda, B + kb, ldb, B + kb + w, ldb, stream);
    }
    long last = ((n - 1) / bs) * bs;
    for (long kb = last; kb >= 0; kb -= bs) {
      int w = (int)std::min<long>(bs, n - kb);
      hipLaunchKernelGGL(k_trsv_block<T>, dim3((unsigned)nrhs), dim3(256), 0, stream, LU, lda, kb, w, B, ldb, 0, 0);
      if (kb > 0) gemm_sub(kb, nrhs, (long)w, LU + kb * lda, lda, B + kb, ldb, B, ldb, stream);
    }
  } else if (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c') {
    int cj = (trans == 'C' || trans == 'c');
    hipLaunchKernelGGL(k_trsv_trans<T>, dim3((unsigned)nrhs), dim3(256), 0, stream, LU, lda, n, B, ldb, 1, cj);
    apply_row_perm(B, ldb, n, nrhs, d_ipiv, false, tmp, stream);
  } else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// log det of an LU-factored matrix: sum of log |U_ii| and of arg U_ii, plus
// pi for every row interchange (deterministic two-pass reduction)
template <typename T>
__global__ void __launch_bounds__(256) k_logdet(const T *LU, long n, long lda, const int *ipiv, double *out)
{
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (long i = threadIdx.x; i < n; i += 256) {
    T d = LU[i + i * lda];
    double re = re_of(d), im = im_of(d);
    a += 0.5 * log(re * re + im * im);
    b += atan2(im, re);
    if (ipiv[i] != i + 1) b += 3.14159265358979323846;
  }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if (threadIdx.x < q) { s0[threadIdx.x] += s0[threadIdx.x + q]; s1[threadIdx.x] += s1[threadIdx.x + q]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

// trace of a square matrix (complex sum of the diagonal)
template <typename T>
__global__ void __launch_bounds__(256) k_trace(const T *A, long n, long lda, double *out)
{
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (long i = threadIdx.x; i < n; i += 256) { T d = A[i + i * lda]; a += re_of(d); b += im_of(d); }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if (threadIdx.x < q) { s0[threadIdx.x] += s0[threadIdx.x + q]; s1[threadIdx.x] += s1[threadIdx.x + q]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

template <typename T>
int lu_logdet(const T *LU, long n, long lda, const int *ipiv, double out[2], hipStream_t s)
{
  double *d = (double *)ws_get(WS_SMALL, 256);
  if (!d) return -2;
  hipLaunchKernelGGL(k_logdet<T>, dim3(1), dim3(256), 0, s, LU, n, lda, ipiv, d);
  hipMemcpyAsync(out, d, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -3;
}

template <typename T>
int mat_trace(const T *A, long n, long lda, double out[2], hipStream_t s)
{
  double *d = (double *)ws_get(WS_SMALL, 256);
  if (!d) return -2;
  hipLaunchKernelGGL(k_trace<T>, dim3(1), dim3(256), 0, s, A, n, lda, d);
  hipMemcpyAsync(out, d, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -3;
}

template int lu_logdet<double2>(const double2 *, long, long, const int *, double *, hipStream_t);
template int lu_logdet<double>(const double *, long, long, const int *, double *, hipStream_t);
template int mat_trace<double>(const double *, long, long, double *, hipStream_t);
template int lu_factor<double>(double *, long, long, int *, int *, hipStream_t, int, int);
template int lu_solve<double>(const double *, long, long, const int *, char, long, double *, long, hipStream_t,
                              SolveWs<double> *);
template int lu_logdet<float2>(const float2 *, long, long, const int *, double *, hipStream_t);
template int mat_trace<double2>(const double2 *, long, long, double *, hipStream_t);
template int mat_trace<float2>(const float2 *, long, long, double *, hipStream_t);
template int lu_factor<double2>(double2 *, long, long, int *, int *, hipStream_t, int, int);
template int lu_factor<float2>(float2 *, long, long, int *, int *, hipStream_t, int, int);
template int lu_solve<double2>(const double2 *, long, long, const int *, char, long, double2 *, long, hipStream_t,
                               SolveWs<double2> *);
template int lu_solve<float2>(const float2 *, long, long, const int *, char, long, float2 *, long, hipStream_t,
                              SolveWs<float2> *);
template size_t solve_ws_bytes<double2>(long);
template size_t solve_ws_bytes<float2>(long);
template SolveWs<double2> solve_ws_carve<double2>(char *, long);
template SolveWs<float2> solve_ws_carve<float2>(char *, long);

// ---------------------------------------------------------------------
// one-right-hand-side solve that allocates nothing: the buffers of a
// SolveWs (W, inverses of the 64x64 diagonal blocks of L and U, composed
// permutation, one-launch words) are carved from a fixed device array,
// one copy per GPU, held until the process ends.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------
constexpr long r256_const(long b) { return 256 * ((b + 255) / 256); }
constexpr long NOALLOC_NBLK = (NOALLOC_MAX_N + TB - 1) / TB;
constexpr long NOALLOC_BYTES = r256_const((long)sizeof(double2) * NOALLOC_MAX_N) +
                               2 * r256_const((long)sizeof(double2) * NOALLOC_NBLK * TB * TB) +
                               r256_const((long)sizeof(int) * NOALLOC_MAX_N) +
                               r256_const((long)sizeof(unsigned long long) * (2 * NOALLOC_NBLK + 8));
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

long noalloc_scratch_bytes() { return NOALLOC_BYTES; }

// device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (the array's name in host code is only a
// host-side stand-in, never a device address)
static void *noalloc_scratch_ptr()
{
  static std::mutex mu;
  static void *ptr[64] = {nullptr};
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[d]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    ptr[d] = p;
  }
  return ptr[d];
}

template <typename T>
int lu_solve_noalloc(const T *LU, long n, long lda, const int *d_ipiv, T *B, hipStream_t stream)
{
  if (n < 0 || lda < std::max(1L, n)) return -1;
  if (n > NOALLOC_MAX_N || (long)solve_ws_bytes<T>(n) > NOALLOC_BYTES) return -7;
  if (n == 0) return 0;
  char *base = (char *)noalloc_scratch_ptr();
  if (!base) return -3;
  SolveWs<T> w = solve_ws_carve<T>(base, n);
  w.inv_ready = false;   // the array holds whatever the last call left
  return lu_solve<T>(LU, n, lda, d_ipiv, 'N', 1, B, std::max(1L, n), stream, &w);
}
template int lu_solve_noalloc<double2>(const double2 *, long, long, const int *, double2 *, hipStream_t);
template int lu_solve_noalloc<float2>(const float2 *, long, long, const int *, float2 *, hipStream_t);

// ---------------------------------------------------------------------
// refinement helpers
// ---------------------------------------------------------------------
#define RES_COLCHUNK 512
__global__ void __launch_bounds__(256)
k_resid_partial(const double2 *A, long n, long lda, const double2 *x, double2 *work, long sa, long sx, long sw)
{
  {
    const long z = blockIdx.z;   // instance of a batch
    A += z * sa; x += z * sx; work += z * sw;
  }
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  long c0 = (long)blockIdx.y * RES_COLCHUNK, c1 = std::min(n, c0 + RES_COLCHUNK);
  if (i >= n) return;
  double sr = 0, si = 0;
  for (long c = c0; c < c1; c++) {
    double2 a = A[i + c * lda], v = x[c];
    sr += a.x * v.x - a.y * v.y;
    si += a.x * v.y + a.y * v.x;
  }
  work[blockIdx.y * n + i] = make_double2(sr, si);
}

__global__ void k_resid_final(long n, int nch, const double2 *b, const double2 *work, double2 *r, long sb, long sw,
                              long srr)
{
  {
    const long z = blockIdx.y;   // instance of a batch
    if (b) b += z * sb;
    work += z * sw; r += z * srr;
  }
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sr = 0, si = 0;
  for (int q = 0; q < nch; q++) { sr += work[q * n + i].x; si += work[q * n + i].y; }
  r[i] = b ? make_double2(b[i].x - sr, b[i].y - si) : make_double2(sr, si);
}

long residual_chunks(long n) { return (n + RES_COLCHUNK - 1) / RES_COLCHUNK; }

int residual_c128(const double2 *A, long n, long lda, const double2 *x, const double2 *b, double2 *r,
                  double2 *work, hipStream_t stream)
{
  int nch = (int)((n + RES_COLCHUNK - 1) / RES_COLCHUNK);
  hipLaunchKernelGGL(k_resid_partial, dim3((unsigned)((n + 255) / 256), nch), dim3(256), 0, stream, A, n, lda, x, work,
                     0L, 0L, 0L);
  hipLaunchKernelGGL(k_resid_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, nch, b, work, r, 0L, 0L,
                     0L);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void k_c128_to_c64(const double2 *x, float2 *y, long rows, long cols, long ldx, long ldy)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i >= rows) return;
  for (; c < cols; c += gridDim.y) {
    double2 v = x[i + c * ldx];
    y[i + c * ldy] = make_float2((float)v.x, (float)v.y);
  }
}

int convert_c128_to_c64(const double2 *x, float2 *y, long rows, long cols, long ldx, long ldy, hipStream_t stream)
{
  unsigned gy = (unsigned)std::min<long>(cols, 65535);
  hipLaunchKernelGGL(k_c128_to_c64, dim3((unsigned)((rows + 255) / 256), gy), dim3(256), 0, stream, x, y, rows, cols, ldx, ldy);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// real parts of a complex128 block into a real matrix (column-major, or
// packed upper storage of a square block: dst[r + c (c+1)/2], r <= c)
__global__ void k_re_part(const double2 *x, long ldx, double *y, long ldy, long rows, long cols, int packed)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long c = blockIdx.y; c < cols; c += gridDim.y) {
    if (i >= rows) return;
    const double v = x[i + c * ldx].x;
    if (!packed) y[i + c * ldy] = v;
    else if (i <= c) y[i + c * (c + 1) / 2] = v;
  }
}

int real_part_c128(const double2 *x, long ldx, double *y, long ldy, long rows, long cols, int packed,
                   hipStream_t stream)
{
  if (rows <= 0 || cols <= 0) return 0;
  unsigned gy = (unsigned)std::min<long>(cols, 65535);
  hipLaunchKernelGGL(k_re_part, dim3((unsigned)((rows + 255) / 256), gy), dim3(256), 0, stream, x, ldx, y, ldy, rows,
                     cols, packed);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// full n x n matrix from packed upper storage (symmetric: A(r,c) = A(c,r))
template <typename T>
__global__ void k_unpack_sym(const T *P, T *A, long lda, long n)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long c = blockIdx.y; c < n; c += gridDim.y) {
    if (r >= n) return;
    const long lo = r < c ? r : c, hi = r < c ? c : r;
    A[r + c * lda] = P[lo + hi * (hi + 1) / 2];
  }
}

template <typename T>
int unpack_symmetric(const T *P, T *A, long lda, long n, hipStream_t stream)
{
  if (n <= 0) return 0;
  unsigned gy = (unsigned)std::min<long>(n, 65535);
  hipLaunchKernelGGL(k_unpack_sym<T>, dim3((unsigned)((n + 255) / 256), gy), dim3(256), 0, stream, P, A, lda, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
template int unpack_symmetric<double>(const double *, double *, long, long, hipStream_t);
template int unpack_symmetric<double2>(const double2 *, double2 *, long, long, hipStream_t);
template int unpack_symmetric<float2>(const float2 *, float2 *, long, long, hipStream_t);

// identity matrix (n x n, column-major)
template <typename T>
__global__ void k_identity(T *A, long lda, long n)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long c = blockIdx.y; c < n; c += gridDim.y) {
    if (r >= n) return;
    A[r + c * lda] = r == c ? unit_v<T>() : zero_v<T>();
  }
}

// first zero diagonal entry (1-based) of U, or 0
template <typename T>
__global__ void k_first_zero_diag(const T *LU, long lda, long n, int *out)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (!nonzero(LU[i + i * lda])) atomicMin(out, (int)(i + 1));
}

// LUInvert (zgetri / dgetri semantics): A <- A^{-1} from its LU factors
// (unit-lower L, U, ipiv), as the solution of A X = I with nrhs = n; W holds
// n x n elements plus one int.  info = first zero diagonal of U (A unchanged)
template <typename T>
int lu_invert(T *A, long n, long lda, const int *ipiv, T *W, int *info, hipStream_t s)
{
  *info = 0;
  if (n <= 0) return 0;
  int *dz = (int *)(W + n * n);
  const int big = 0x7F7F7F7F;          // every byte 0x7F: larger than any index
  hipMemsetAsync(dz, 0x7F, sizeof(int), s);
  hipLaunchKernelGGL(k_first_zero_diag<T>, dim3((unsigned)std::min<long>(1024, (n + 255) / 256)), dim3(256), 0, s, A,
                     lda, n, dz);
  int z = 0;
  hipMemcpyAsync(&z, dz, sizeof(int), hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) return -3;
  if (z != big) { *info = z; return 0; }
  unsigned gy = (unsigned)std::min<long>(n, 65535);
  hipLaunchKernelGGL(k_identity<T>, dim3((unsigned)((n + 255) / 256), gy), dim3(256), 0, s, W, n, n);
  int rc = lu_solve<T>(A, n, lda, ipiv, 'N', n, W, n, s);
  if (rc) return rc;
  hipMemcpy2DAsync(A, sizeof(T) * lda, W, sizeof(T) * n, sizeof(T) * n, n, hipMemcpyDeviceToDevice, s);
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -3;
}
template int lu_invert<double2>(double2 *, long, long, const int *, double2 *, int *, hipStream_t);
template int lu_invert<float2>(float2 *, long, long, const int *, float2 *, int *, hipStream_t);
template int lu_invert<double>(double *, long, long, const int *, double *, int *, hipStream_t);

__global__ void k_c64_to_c128(const float2 *x, double2 *y, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = make_double2(x[i].x, x[i].y);
}
int convert_c64_to_c128(const float2 *x, double2 *y, long n, hipStream_t stream)
{
  hipLaunchKernelGGL(k_c64_to_c128, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x, y, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void k_axpy64(const float2 *d, double2 *x, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { x[i].x += (double)d[i].x; x[i].y += (double)d[i].y; }
}
int axpy_c64_into_c128(const float2 *d, double2 *x, long n, hipStream_t stream)
{
  hipLaunchKernelGGL(k_axpy64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, d, x, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void k_absmax(const double2 *x, long n, double *out)
{
  __shared__ double s[256];
  double m = 0;
  for (long i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(x[i].x) + fabs(x[i].y));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) { if (threadIdx.x < q) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + q]); __syncthreads(); }
  if (threadIdx.x == 0) *out = s[0];
}
double absmax_c128(const double2 *x, long n, double *work, hipStream_t stream)
{
  hipLaunchKernelGGL(k_absmax, dim3(1), dim3(256), 0, stream, x, n, work);
  double h = 0;
  hipMemcpyAsync(&h, work, sizeof(double), hipMemcpyDeviceToHost, stream);
  hipStreamSynchronize(stream);
  return h;
}

// ---------------------------------------------------------------------
// Batched complex64 LU with tournament pivoting: nbat matrices of equal
// order n at A + b sa, factored in lockstep, every step one launch over
// the whole batch (panels: (G + 1) x nbat workgroups; updates: grid.y or
// grid.z = instance).  No lookahead (the batch fills the GPU), no host
// synchronization, no allocation and no memset, so that a stream capture
// records it.  The per-instance panel scratch (layout of lu_factor's
// complex64 tournament path) is cleared by a kernel at the start.
// ---------------------------------------------------------------------
struct BatchLayout {
  size_t off_info, off_plan, off_li, off_ts, total;
  long Gmax, npanels;
  size_t off_cv, off_ci, off_cr, off_rc;
};

static BatchLayout batch_layout(long n)
{
  auto r256 = [](size_t b) { return 256 * ((b + 255) / 256); };
  BatchLayout L;
  L.Gmax = (n + 255) / 256;
  L.npanels = (n + PW - 1) / PW;
  L.off_info = 96 * sizeof(unsigned);
  L.off_plan = L.off_info + 64;
  L.off_cv = L.off_plan + r256(sizeof(int) * PLAN_LEN * L.npanels);
  L.off_ci = L.off_cv + r256(sizeof(double) * 2 * L.Gmax);
  L.off_cr = L.off_ci + r256(sizeof(unsigned long long) * 2 * L.Gmax);
  L.off_rc = L.off_cr + r256(sizeof(float2) * 2 * L.Gmax * PW);
  L.off_li = L.off_rc + r256(sizeof(float2) * 2 * PW);
  L.off_ts = L.off_li + r256(sizeof(float2) * PW * PW * L.npanels);
  const size_t ts_words = (size_t)ts_root_off(ts_levels(L.Gmax), L.Gmax) + TS_ROOT;
  L.total = r256(L.off_ts + sizeof(unsigned long long) * ts_words);
  return L;
}

size_t lu_batch_scratch_bytes(long n) { return batch_layout(n).total; }

__global__ void k_batch_status(const char *scratch, long sS, long off_info, int nbat, int *status)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbat) return;
  const char *base = scratch + (long)b * sS;
  status[2 * b] = *(const int *)(base + off_info);
  status[2 * b + 1] = (int)((const unsigned *)base)[64];
}

int lu_batch_max(long n)
{
  int dev = 0, ncu = 0, perCU = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void *)k_panel_ts, 256, 0);
  const long per = (n + 255) / 256 + 1;
  return (int)((long)std::max(1, ncu) * std::max(1, perCU) / per);
}

int lu_factor_batch(float2 *A, long n, long lda, int nbat, long sa, int *ipiv, long sp, void *scratch, int *status,
                    hipStream_t s, int nb)
{
  if (n <= 0 || nbat <= 0 || lda < n || !scratch || !status) return -1;
  if (nb <= 0 || nb > 256 || nb % PW) return -1;
  const BatchLayout L = batch_layout(n);
  if (L.Gmax > 64 * KEYS_PER_LANE) return -5;
  if (nbat > lu_batch_max(n)) return -5;      // every panel workgroup of the batch must be co-resident
  const long sS = (long)L.total;
  char *base = (char *)scratch;
  const long nwords = (long)nbat * sS / 8;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s,
                     (unsigned long long *)base, nwords);
  PanelScratch S;
  S.bar = (unsigned *)base;
  int *dinfo = (int *)(base + L.off_info);
  int *plans = (int *)(base + L.off_plan);
  S.candv = (double *)(base + L.off_cv);
  S.ckey = (unsigned long long *)(base + L.off_ci);
  S.candrow = (void *)(base + L.off_cr);
  S.rowc = (void *)(base + L.off_rc);
  S.linv = (void *)(base + L.off_li);
  S.ts = (unsigned long long *)(base + L.off_ts);
  const float2 *linv = (const float2 *)S.linv;
  const bool g3m = mixed_gemm_3m();
  BStride bs;
  bs.a = sa; bs.p = sp; bs.s = sS;
  int gmaxi = (int)L.Gmax;
  for (long K = 0; K < n; K += nb) {
    const long jb = std::min<long>(nb, n - K);
    const long p0 = K / PW;
    for (long j0 = K; j0 < K + jb; j0 += PW) {
      int w = (int)std::min<long>(PW, K + jb - j0);
      long m = n - j0, G = (m + 255) / 256, r0 = j0;
      S.plan = plans + (j0 / PW) * PLAN_LEN;
      unsigned seq0 = (unsigned)(j0 + 1);
      void *args[] = {(void *)&A, (void *)&lda, (void *)&r0, (void *)&m, (void *)&w, (void *)&ipiv,
                      (void *)&dinfo, (void *)&S, (void *)&gmaxi, (void *)&seq0, (void *)&bs};
      int tok = prof_begin(s, PK_PANEL_C64);
      hipError_t e = hipLaunchKernel((const void *)k_panel_ts, dim3((unsigned)G + 1, (unsigned)nbat), dim3(256), args,
                                     0, s);
      prof_end(s, tok, 0.0);
      if (e != hipSuccess) return -3;
      swap_trmm(A, lda, K, K + jb, j0, j0 + w, S.plan, linv + (j0 / PW) * PW * PW, w, j0, s, nbat, bs);
      const long nr = K + jb - (j0 + w);
      if (nr > 0)
        gemm_sub(n - j0 - w, nr, (long)w, A + (j0 + w) + j0 * lda, lda, A + j0 + (j0 + w) * lda, lda,
                 A + (j0 + w) + (j0 + w) * lda, lda, s, 0, nbat, sa, g3m);
    }
    const long c1 = K + jb;
    if (c1 < n) {
      const int bplans = (int)((jb + PW - 1) / PW);
      laswp_plans(A, lda, c1, n, 0, 0, plans + p0 * PLAN_LEN, bplans, s, nbat, bs);
      trsm_blk(A, lda, K, (int)jb, linv + p0 * PW * PW, A + K + c1 * lda, lda, n - c1, s, nbat, bs);
      gemm_sub(n - c1, n - c1, jb, A + c1 + K * lda, lda, A + K + c1 * lda, lda, A + c1 + c1 * lda, lda, s, 0, nbat,
               sa, g3m);
    }
  }
  // interchanges of each block on the columns left of it
  for (long K = nb; K < n; K += nb) {
    const long jb = std::min<long>(nb, n - K);
    laswp_plans(A, lda, 0, K, 0, 0, plans + (K / PW) * PLAN_LEN, (int)((jb + PW - 1) / PW), s, nbat, bs);
  }
  hipLaunchKernelGGL(k_batch_status, dim3((unsigned)((nbat + 63) / 64)), dim3(64), 0, s, (const char *)base, sS,
                     (long)L.off_info, nbat, status);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Batched one-right-hand-side solves with batched complex64 factors:
// prepare (composed row permutations and inverses of the 64 x 64 diagonal
// blocks of L and U, once per factorization), then any number of solves
// of B (n per instance, sb apart) in place.  No host synchronization.
struct BatchSolveLayout {
  size_t off_invL, off_invU, off_W, total;
};
static BatchSolveLayout batch_solve_layout(long n)
{
  auto r256 = [](size_t b) { return 256 * ((b + 255) / 256); };
  const long nblk = (n + TB - 1) / TB;
  BatchSolveLayout L;
  L.off_invL = r256(sizeof(int) * n);
  L.off_invU = L.off_invL + r256(sizeof(float2) * nblk * TB * TB);
  L.off_W = L.off_invU + r256(sizeof(float2) * nblk * TB * TB);
  L.total = L.off_W + r256(sizeof(float2) * n);
  return L;
}
size_t lu_solve_batch_bytes(long n) { return batch_solve_layout(n).total; }

int lu_solve_batch_prepare(const float2 *LU, long n, long lda, int nbat, long sa, const int *ipiv, long sp, void *ws,
                           hipStream_t s)
{
  if (n <= 0 || nbat <= 0 || n > PERM_LDS_MAX || !ws) return -1;
  const BatchSolveLayout L = batch_solve_layout(n);
  const long sw = (long)L.total, nblk = (n + TB - 1) / TB;
  char *base = (char *)ws;
  hipLaunchKernelGGL(k_perm_build, dim3((unsigned)nbat), dim3(64), 0, s, ipiv, n, (int *)base, sp, sw / 4);
  hipLaunchKernelGGL((k_diag_inv<float2, true>), dim3((unsigned)nblk, 1, (unsigned)nbat), dim3(TB), 0, s, LU, lda, n,
                     (float2 *)(base + L.off_invL), sa, sw / 8);
  hipLaunchKernelGGL((k_diag_inv<float2, false>), dim3((unsigned)nblk, 1, (unsigned)nbat), dim3(TB), 0, s, LU, lda, n,
                     (float2 *)(base + L.off_invU), sa, sw / 8);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void k_copy_batch(const float2 *in, long sin, float2 *out, long sout, long n)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x, z = blockIdx.z;
  if (i < n) out[z * sout + i] = in[z * sin + i];
}

int lu_solve_batch(const float2 *LU, long n, long lda, int nbat, long sa, float2 *B, long sb, void *ws, hipStream_t s)
{
  if (n <= 0 || nbat <= 0 || !ws) return -1;
  const BatchSolveLayout L = batch_solve_layout(n);
  const long sw = (long)L.total;
  char *base = (char *)ws;
  const int *perm = (const int *)base;
  const float2 *invL = (const float2 *)(base + L.off_invL), *invU = (const float2 *)(base + L.off_invU);
  float2 *W = (float2 *)(base + L.off_W);
  const long swp = sw / 4, swe = sw / 8;   // the stride in ints and in complex64 elements
  // W = P B
  hipLaunchKernelGGL((k_perm_apply<float2, true>), dim3((unsigned)((n + 255) / 256), 1, (unsigned)nbat), dim3(256), 0,
                     s, (const float2 *)B, n, W, n, perm, n, sb, swe, swp);
  // L y = W: W is the running right-hand side, y lands in B
  for (long kb = 0; kb < n; kb += TB) {
    const int bs = (int)std::min<long>(TB, n - kb);
    const long rows = n - kb - bs;
    dim3 grid((unsigned)std::max<long>(1, (rows + 255) / 256), 1, (unsigned)nbat);
    hipLaunchKernelGGL((k_trsv_inv<float2, true>), grid, dim3(256), 0, s, LU, lda, n, kb, bs,
                       invL + (kb / TB) * TB * TB, W, n, B, n, sa, swe, swe, sb);
  }
  // U x = y: B is the running right-hand side, x lands in W
  const long last = ((n - 1) / TB) * TB;
  for (long kb = last; kb >= 0; kb -= TB) {
    const int bs = (int)std::min<long>(TB, n - kb);
    dim3 grid((unsigned)std::max<long>(1, (kb + 255) / 256), 1, (unsigned)nbat);
    hipLaunchKernelGGL((k_trsv_inv<float2, false>), grid, dim3(256), 0, s, LU, lda, n, kb, bs,
                       invU + (kb / TB) * TB * TB, B, n, W, n, sa, swe, sb, swe);
  }
  // B = x
  hipLaunchKernelGGL(k_copy_batch, dim3((unsigned)((n + 255) / 256), 1, (unsigned)nbat), dim3(256), 0, s,
                     (const float2 *)W, swe, B, sb, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// per instance of a batch: max|r_i| / max|b_i| (|re| + |im|), vectors n apart
__global__ void k_batch_relres(const double2 *r, const double2 *b, long n, double *out)
{
  __shared__ double sr[256], sb[256];
  const long z = blockIdx.x;
  r += z * n; b += z * n;
  double mr = 0, mb = 0;
  for (long i = threadIdx.x; i < n; i += 256) {
    mr = fmax(mr, fabs(r[i].x) + fabs(r[i].y));
    mb = fmax(mb, fabs(b[i].x) + fabs(b[i].y));
  }
  sr[threadIdx.x] = mr; sb[threadIdx.x] = mb;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if (threadIdx.x < q) { sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + q]); sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + q]); }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[z] = sb[0] > 0 ? sr[0] / sb[0] : sr[0];
}
int batch_relres(const double2 *r, const double2 *b, long n, int nbat, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_relres, dim3((unsigned)nbat), dim3(256), 0, s, r, b, n, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void k_zero_c128(double2 *x, long n)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = make_double2(0.0, 0.0);
}
int zero_c128(double2 *x, long n, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_zero_c128, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// r = b - A x for a batch (A + z sa, x / b / r n apart, work n * chunks apart)
int residual_c128_batch(const double2 *A, long n, long lda, long sa, int nbat, const double2 *x, const double2 *b,
                        double2 *r, double2 *work, hipStream_t stream)
{
  const int nch = (int)((n + RES_COLCHUNK - 1) / RES_COLCHUNK);
  hipLaunchKernelGGL(k_resid_partial, dim3((unsigned)((n + 255) / 256), nch, (unsigned)nbat), dim3(256), 0, stream, A,
                     n, lda, x, work, sa, n, n * (long)nch);
  hipLaunchKernelGGL(k_resid_final, dim3((unsigned)((n + 255) / 256), (unsigned)nbat), dim3(256), 0, stream, n, nch, b,
                     work, r, n, n * (long)nch, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace scuffx
