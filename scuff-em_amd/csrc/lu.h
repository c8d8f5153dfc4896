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

// T = double2 (complex128) or float2 (complex64).  fast = 1 solves the
// block rows with the explicit inverses of the 32x32 unit-lower panel
// blocks (products instead of substitution): same pivots, slightly larger
// rounding; meant for preconditioner factorizations.
template <typename T>
int lu_factor(T *A, long n, long lda, int *d_ipiv, int *info, hipStream_t stream, int nb = 256,
              int fast = 0);

// solves op(A) X = B with the factors of lu_factor; trans 'N', 'T', 'C'
template <typename T>
int lu_solve(const T *LU, long n, long lda, const int *d_ipiv, char trans, long nrhs, T *B,
             long ldb, hipStream_t stream);

// log |det| and arg det of an LU-factored matrix (out[0], out[1]; host)
template <typename T>
int lu_logdet(const T *LU, long n, long lda, const int *ipiv, double out[2], hipStream_t stream);
// complex trace of a square matrix (out[0] + i out[1]; host)
template <typename T>
int mat_trace(const T *A, long n, long lda, double out[2], hipStream_t stream);

// r = b - A x (complex128, A column-major n x n); all device pointers
int residual_c128(const double2 *A, long n, long lda, const double2 *x, const double2 *b,
                  double2 *r, double2 *work, hipStream_t stream);
// y = (complex64) x and back
int convert_c128_to_c64(const double2 *x, float2 *y, long rows, long cols, long ldx, long ldy,
                        hipStream_t stream);
int convert_c64_to_c128(const float2 *x, double2 *y, long n, hipStream_t stream);
// x += (complex128) d
int axpy_c64_into_c128(const float2 *d, double2 *x, long n, hipStream_t stream);
// max_i |re|+|im| of a vector (host result)
double absmax_c128(const double2 *x, long n, double *work, hipStream_t stream);

}  // namespace scuffx
