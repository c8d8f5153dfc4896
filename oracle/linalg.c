/*
 * linalg.c -- oracle: LU factorization and solve with the semantics of
 * LAPACK zgetrf/zgetrs as called by HMatrix::LUFactorize / LUSolve
 * (libs/libhmat/LBWrappers.cc:151-245).  TEST INFRASTRUCTURE ONLY.
 *
 * Partial pivoting on |re|+|im| (izamax), first maximal index wins;
 * unit-lower L stored below the diagonal, U on and above; ipiv is
 * 1-based; info = 1-based index of the first exactly-zero pivot.
 * The column update order is the right-looking unblocked one (zgetf2).
 * LAPACK itself is not vendored in the reference (lapack.h declares
 * prototypes only), so this restates the published LAPACK algorithm.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define A_(i, j) A[(size_t)(i) + (size_t)(j)*(size_t)lda]

int o_zgetrf(int n, zdouble *A, int lda, int *ipiv)
{
  int info = 0;
  for (int j = 0; j < n; j++) {
    int p = j;
    double best = -1.0;
    for (int i = j; i < n; i++) {
      double v = fabs(creal(A_(i, j))) + fabs(cimag(A_(i, j)));
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p + 1;
    if (A_(p, j) != 0.0) {
      if (p != j)
        for (int c = 0; c < n; c++) { zdouble t = A_(j, c); A_(j, c) = A_(p, c); A_(p, c) = t; }
      zdouble r = 1.0/A_(j, j);
      for (int i = j+1; i < n; i++) A_(i, j) *= r;
    } else if (info == 0) info = j + 1;
    for (int c = j+1; c < n; c++) {
      zdouble u = A_(j, c);
      if (u != 0.0)
        for (int i = j+1; i < n; i++) A_(i, c) -= A_(i, j)*u;
    }
  }
  return info;
}

int o_zgetrs(char trans, int n, int nrhs, const zdouble *A, int lda,
             const int *ipiv, zdouble *B, int ldb)
{
  if (n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1)) return -1;
  for (int r = 0; r < nrhs; r++) {
    zdouble *b = B + (size_t)r*ldb;
    if (trans == 'N' || trans == 'n') {
      for (int i = 0; i < n; i++) { int p = ipiv[i] - 1; if (p != i) { zdouble t = b[i]; b[i] = b[p]; b[p] = t; } }
      for (int j = 0; j < n; j++) { zdouble x = b[j]; if (x != 0.0) for (int i = j+1; i < n; i++) b[i] -= x*A_(i, j); }
      for (int j = n-1; j >= 0; j--) { if (b[j] != 0.0) { b[j] /= A_(j, j); zdouble x = b[j]; for (int i = 0; i < j; i++) b[i] -= x*A_(i, j); } }
    } else {
      int cj = (trans == 'C' || trans == 'c');
      for (int j = 0; j < n; j++) {
        zdouble s = b[j];
        for (int i = 0; i < j; i++) s -= (cj ? conj(A_(i, j)) : A_(i, j))*b[i];
        b[j] = s/(cj ? conj(A_(j, j)) : A_(j, j));
      }
      for (int j = n-1; j >= 0; j--) {
        zdouble s = b[j];
        for (int i = j+1; i < n; i++) s -= (cj ? conj(A_(i, j)) : A_(i, j))*b[i];
        b[j] = s;
      }
      for (int i = n-1; i >= 0; i--) { int p = ipiv[i] - 1; if (p != i) { zdouble t = b[i]; b[i] = b[p]; b[p] = t; } }
    }
  }
  return 0;
}
