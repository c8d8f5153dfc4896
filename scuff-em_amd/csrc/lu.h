// lu.h -- dense complex LU factorization and solves on one GPU.
//
// Semantics of LAPACK zgetrf/zgetrs as used by HMatrix::LUFactorize and
// HMatrix::LUSolve (libs/libhmat/LBWrappers.cc:151-245): partial pivoting
// on |re|+|im| (first maximal row wins), unit-lower L below the diagonal,
// U on and above it, 1-based row interchanges ipiv, info = first zero
// pivot (1-based) or 0.  Matrices are column-major, interleaved complex.
#pragma once

#include <hip/hip_runtime.h>

namespace scuffx {

// T = double2 (complex128), float2 (complex64) or double (real, the LHM_REAL
// matrices of imaginary frequencies: dgetrf/dgetrs semantics).  fast = 1 solves the
// block rows with the explicit inverses of the 32x32 unit-lower panel
// blocks (products instead of substitution), slightly larger rounding; for
// complex64 it also pivots by tournament and updates with the 3M product
// (see lu.hip); meant for preconditioner factorizations.
template <typename T>
int lu_factor(T *A, long n, long lda, int *d_ipiv, int *info, hipStream_t stream, int nb = 256,
              int fast = 0);

// Caller-provided buffers of a one-right-hand-side solve (trans 'N'):
// the intermediate vector W (n), the inverses of the 64x64 diagonal blocks
// of L and U (nblk * 64 * 64 each, nblk = ceil(n / 64)), the composed row
// permutation (n ints) and the one-launch solve's words (2 nblk + 8).
// inv_ready: invL / invU and perm already hold the inverses and the
// composed row permutation for these factors (set by the first solve;
// clear it for new factors).
template <typename T>
struct SolveWs {
  T *W = nullptr, *invL = nullptr, *invU = nullptr;
  int *perm = nullptr;
  unsigned long long *chain = nullptr;
  bool inv_ready = false;
};
// bytes of the buffers above for n rows, each part rounded to 256 bytes
template <typename T>
size_t solve_ws_bytes(long n);
// carves the buffers from `base` (solve_ws_bytes(n) bytes)
template <typename T>
SolveWs<T> solve_ws_carve(char *base, long n);

// solves op(A) X = B with the factors of lu_factor; trans 'N', 'T', 'C'.
// With ws (trans 'N', nrhs 1 only) it allocates nothing.
template <typename T>
int lu_solve(const T *LU, long n, long lda, const int *d_ipiv, char trans, long nrhs, T *B,
             long ldb, hipStream_t stream, SolveWs<T> *ws = nullptr);

// lu_solve(trans 'N', one right-hand side) with its buffers in a fixed
// device array (no allocation; n <= NOALLOC_MAX_N, else -7).  Calls that
// overlap on one GPU race: they share that GPU's array.
constexpr long NOALLOC_MAX_N = 16384;
template <typename T>
int lu_solve_noalloc(const T *LU, long n, long lda, const int *d_ipiv, T *B, hipStream_t stream);
long noalloc_scratch_bytes();

// Batched complex64 LU with tournament pivoting of nbat matrices of order
// n at A + b sa (ipiv + b sp), all in lockstep; no host synchronization,
// allocation or memset (a stream capture can record it).  scratch:
// nbat * lu_batch_scratch_bytes(n) bytes; status (device): per instance
// {LAPACK info, bounded-wait time-out flag}.  Returns -5 if the batch's
// panel workgroups cannot all be resident (nbat > lu_batch_max(n)).
size_t lu_batch_scratch_bytes(long n);
int lu_batch_max(long n);
// nb: columns per block (a multiple of 32, at most 256)
int lu_factor_batch(float2 *A, long n, long lda, int nbat, long sa, int *ipiv, long sp, void *scratch, int *status,
                    hipStream_t s, int nb = 256);
// one-right-hand-side solves with those factors: prepare once per
// factorization (permutations, inverses of the diagonal blocks into ws,
// nbat * lu_solve_batch_bytes(n) bytes), then solve B (n per instance, sb
// apart) in place as often as needed
size_t lu_solve_batch_bytes(long n);
int lu_solve_batch_prepare(const float2 *LU, long n, long lda, int nbat, long sa, const int *ipiv, long sp, void *ws,
                           hipStream_t s);
int lu_solve_batch(const float2 *LU, long n, long lda, int nbat, long sa, float2 *B, long sb, void *ws, hipStream_t s);
// r = b - A x for a batch (A + z sa; x, b, r n apart; work n * chunks apart)
// per instance: max|r| / max|b| over n entries (|re| + |im|), out on the device
int batch_relres(const double2 *r, const double2 *b, long n, int nbat, double *out, hipStream_t s);
// x[0..n) = 0 by a kernel (no memset node in a captured sequence)
int zero_c128(double2 *x, long n, hipStream_t s);
int residual_c128_batch(const double2 *A, long n, long lda, long sa, int nbat, const double2 *x, const double2 *b,
                        double2 *r, double2 *work, hipStream_t stream);

// log |det| and arg det of an LU-factored matrix (out[0], out[1]; host)
template <typename T>
int lu_logdet(const T *LU, long n, long lda, const int *ipiv, double out[2], hipStream_t stream);
// complex trace of a square matrix (out[0] + i out[1]; host)
template <typename T>
int mat_trace(const T *A, long n, long lda, double out[2], hipStream_t stream);

// r = b - A x (complex128, A column-major n x n); all device pointers;
// b == nullptr gives r = A x.  work: n * residual_chunks(n) elements
int residual_c128(const double2 *A, long n, long lda, const double2 *x, const double2 *b,
                  double2 *r, double2 *work, hipStream_t stream);
long residual_chunks(long n);
// y = (complex64) x and back
int convert_c128_to_c64(const double2 *x, float2 *y, long rows, long cols, long ldx, long ldy,
                        hipStream_t stream);
int convert_c64_to_c128(const float2 *x, double2 *y, long n, hipStream_t stream);
// x += (complex128) d
int axpy_c64_into_c128(const float2 *d, double2 *x, long n, hipStream_t stream);
// real parts of a complex128 block into a real matrix (packed: upper
// storage dst[r + c (c+1)/2] of a square block, r <= c)
int real_part_c128(const double2 *x, long ldx, double *y, long ldy, long rows, long cols, int packed,
                   hipStream_t stream);
// full symmetric n x n matrix from packed upper storage
template <typename T>
int unpack_symmetric(const T *P, T *A, long lda, long n, hipStream_t stream);
// A <- A^{-1} from its LU factors (LUInvert); W: n*n elements + one int;
// *info = first zero diagonal entry of U (A then unchanged), else 0
template <typename T>
int lu_invert(T *A, long n, long lda, const int *ipiv, T *W, int *info, hipStream_t stream);
// max_i |re|+|im| of a vector (host result)
double absmax_c128(const double2 *x, long n, double *work, hipStream_t stream);

}  // namespace scuffx
