// devmath.h -- small host/device numeric helpers for the HIP kernels:
// complex arithmetic, 3-vectors, and the degree-4 triangle rule.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace scuffx {

struct cplx {
  double r, i;
};

__host__ __device__ inline cplx mk(double r, double i) { cplx z; z.r = r; z.i = i; return z; }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return mk(a.r + b.r, a.i + b.i); }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return mk(a.r - b.r, a.i - b.i); }
__host__ __device__ inline cplx operator-(cplx a) { return mk(-a.r, -a.i); }
__host__ __device__ inline cplx operator*(cplx a, cplx b) { return mk(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__host__ __device__ inline cplx operator*(double s, cplx a) { return mk(s * a.r, s * a.i); }
__host__ __device__ inline cplx operator*(cplx a, double s) { return mk(s * a.r, s * a.i); }
__host__ __device__ inline cplx operator/(cplx a, double s) { return mk(a.r / s, a.i / s); }
__host__ __device__ inline cplx& operator+=(cplx &a, cplx b) { a.r += b.r; a.i += b.i; return a; }
__host__ __device__ inline cplx& operator-=(cplx &a, cplx b) { a.r -= b.r; a.i -= b.i; return a; }
__host__ __device__ inline cplx& operator*=(cplx &a, cplx b) { a = a * b; return a; }
__host__ __device__ inline cplx& operator*=(cplx &a, double s) { a.r *= s; a.i *= s; return a; }
__host__ __device__ inline double norm2(cplx a) { return a.r * a.r + a.i * a.i; }
__host__ __device__ inline double cabs_(cplx a) { return hypot(a.r, a.i); }
// Smith's algorithm
__host__ __device__ inline cplx operator/(cplx a, cplx b)
{
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    return mk((a.r + a.i * t) / d, (a.i - a.r * t) / d);
  }
  double t = b.r / b.i, d = b.r * t + b.i;
  return mk((a.r * t + a.i) / d, (a.i * t - a.r) / d);
}
__host__ __device__ inline cplx operator/(double s, cplx b) { return mk(s, 0.0) / b; }
__host__ __device__ inline cplx cexp_(cplx z)
{
  double e = exp(z.r), s, c;
  sincos(z.i, &s, &c);
  return mk(e * c, e * s);
}
__host__ __device__ inline cplx IMUL(cplx z) { return mk(-z.i, z.r); }   // i*z

__host__ __device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__host__ __device__ inline void sub3(const double *a, const double *b, double *c) { c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2]; }
__host__ __device__ inline void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Dunavant degree-4 six-point rule at the 13-digit precision of the
// reference table (libs/libTriInt/TCR.cc:83-91); (u, v, w), sum w = 1/2
#define SCUFFX_TCR4_INIT {                                     \
    0.0915762135098, 0.0915762135098, 0.0549758718276,         \
    0.8168475729805, 0.0915762135098, 0.0549758718276,         \
    0.0915762135098, 0.8168475729805, 0.0549758718276,         \
    0.1081030181681, 0.4459484909160, 0.1116907948390,         \
    0.4459484909160, 0.1081030181681, 0.1116907948390,         \
    0.4459484909160, 0.4459484909160, 0.1116907948390 }

}  // namespace scuffx
