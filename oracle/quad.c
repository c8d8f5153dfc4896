/*
 * quad.c -- oracle: fixed triangle rules and the two adaptive cubature
 * schemes of libSGJC.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  o_tcr          restates GetTCR for orders 4 and 20 (libs/libTriInt/TCR.cc:83-91,
 *                 :346-427, :556-576); the rule tables are the published
 *                 Dunavant degree-4 rule and the 78-point degree-20 rule at
 *                 the 13-digit precision the reference stores.
 *  o_hcubature    restates the serial h-adaptive path of hcubature()
 *                 (libs/libSGJC/hcubature.c:455-615 Genz-Malik 7/5 rule,
 *                 :619-758 Gauss-Kronrod 15, :772-846 heap, :887-931 driver,
 *                 converged.h:10-14 ERROR_INDIVIDUAL).
 *  o_pcubature    restates the p-adaptive Clenshaw-Curtis scheme
 *                 (libs/libSGJC/pcubature.c:60-430).  The Clenshaw-Curtis
 *                 node/weight tables (clencurt.h, absent from the
 *                 reference tree) are regenerated here from the closed
 *                 form in long double, in the layout pcubature.c indexes.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------ */
/* triangle cubature rules: (u, v, w) triples, weights sum to 1/2      */
/* ------------------------------------------------------------------ */
static const double tcr4[18] = {
  0.0915762135098, 0.0915762135098, 0.0549758718276,
  0.8168475729805, 0.0915762135098, 0.0549758718276,
  0.0915762135098, 0.8168475729805, 0.0549758718276,
  0.1081030181681, 0.4459484909160, 0.1116907948390,
  0.4459484909160, 0.1081030181681, 0.1116907948390,
  0.4459484909160, 0.4459484909160, 0.1116907948390 };

#include "tcr20.inc"

const double *o_tcr(int order, int *npts)
{
  if (order == 4) { *npts = 6; return tcr4; }
  if (order == 20) { *npts = 78; return o_tcr20; }
  fprintf(stderr, "oracle: unsupported triangle rule order %d\n", order);
  abort();
}

/* ------------------------------------------------------------------ */
/* h-adaptive cubature                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  double c[3], h[3];
  double vol;
  unsigned split;
  double errmax;
  double *ee;     /* val[j]=ee[2j], err[j]=ee[2j+1] */
} oreg;

typedef struct {
  unsigned dim, fdim, npts;
  double *pts, *vals;
  o_integrand f;
  void *fdata;
} orule;

static double region_vol(const oreg *R, unsigned dim)
{ double v = 1; for (unsigned i = 0; i < dim; i++) v *= 2*R->h[i]; return v; }

/* Genz-Malik degree 7/5, dim 2 or 3 (hcubature.c:455-615) */
static void gm_eval(orule *r, oreg *R)
{
  const double lambda2 = 0.3585685828003180919906451539079374954541;
  const double lambda4 = 0.9486832980505137995996680633298155601160;
  const double lambda5 = 0.6882472016116852977216287342936235251269;
  const double weight2 = 980. / 6561.;
  const double weight4 = 200. / 19683.;
  const double weightE2 = 245. / 486.;
  const double weightE4 = 25. / 729.;
  const double ratio = (lambda2 * lambda2) / (lambda4 * lambda4);
  unsigned dim = r->dim, fdim = r->fdim;
  double weight1 = (double)(12824 - 9120*(int)dim + 400*(int)(dim*dim)) / 19683.;
  double weight3 = (double)(1820 - 400*(int)dim) / 19683.;
  double weight5 = 6859. / 19683. / (double)(1U << dim);
  double weightE1 = (double)(729 - 950*(int)dim + 50*(int)(dim*dim)) / 729.;
  double weightE3 = (double)(265 - 100*(int)dim) / 1458.;
  double p[3], l2[3], l4[3], l5[3];
  double *pts = r->pts;
  unsigned n = 0;
  for (unsigned i = 0; i < dim; i++) { p[i] = R->c[i]; l2[i] = R->h[i]*lambda2; l4[i] = R->h[i]*lambda4; l5[i] = R->h[i]*lambda5; }
#define EMIT() do { memcpy(pts + n*dim, p, dim*sizeof(double)); n++; } while (0)
  EMIT();
  for (unsigned i = 0; i < dim; i++) {
    p[i] = R->c[i] - l2[i]; EMIT();
    p[i] = R->c[i] + l2[i]; EMIT();
    p[i] = R->c[i] - l4[i]; EMIT();
    p[i] = R->c[i] + l4[i]; EMIT();
    p[i] = R->c[i];
  }
  for (unsigned i = 0; i + 1 < dim; i++) {
    p[i] = R->c[i] - l4[i];
    for (unsigned j = i+1; j < dim; j++) {
      p[j] = R->c[j] - l4[j]; EMIT();
      p[i] = R->c[i] + l4[i]; EMIT();
      p[j] = R->c[j] + l4[j]; EMIT();
      p[i] = R->c[i] - l4[i]; EMIT();
      p[j] = R->c[j];
    }
    p[i] = R->c[i];
  }
  /* 2^dim corners in Gray-code order */
  {
    unsigned signs = 0;
    for (unsigned i = 0; i < dim; i++) p[i] = R->c[i] + l5[i];
    for (unsigned i = 0;; ++i) {
      EMIT();
      unsigned d = (unsigned)__builtin_ctz(~i);
      if (d >= dim) break;
      unsigned mask = 1U << d;
      signs ^= mask;
      p[d] = (signs & mask) ? R->c[d] - l5[d] : R->c[d] + l5[d];
    }
  }
#undef EMIT
  for (unsigned k = 0; k < n; k++) r->f(dim, pts + k*dim, r->fdata, fdim, r->vals + k*fdim);

  double diff[3] = {0, 0, 0};
  unsigned nRR = 2*dim*(dim-1), nRR5 = 1U << dim;
  for (unsigned j = 0; j < fdim; j++) {
#define VALS(i) r->vals[(i)*fdim + j]
    double val0 = VALS(0), sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
    unsigned k, k0 = 1;
    for (k = 0; k < dim; k++) {
      double v0 = VALS(k0 + 4*k), v1 = VALS(k0 + 4*k + 1);
      double v2 = VALS(k0 + 4*k + 2), v3 = VALS(k0 + 4*k + 3);
      sum2 += v0 + v1;
      sum3 += v2 + v3;
      diff[k] += fabs(v0 + v1 - 2*val0 - ratio*(v2 + v3 - 2*val0));
    }
    k0 += 4*k;
    for (k = 0; k < nRR; k++) sum4 += VALS(k0 + k);
    k0 += k;
    for (k = 0; k < nRR5; k++) sum5 += VALS(k0 + k);
#undef VALS
    double result = R->vol * (weight1*val0 + weight2*sum2 + weight3*sum3 + weight4*sum4 + weight5*sum5);
    double res5th = R->vol * (weightE1*val0 + weightE2*sum2 + weightE3*sum3 + weightE4*sum4);
    R->ee[2*j] = result;
    R->ee[2*j+1] = fabs(res5th - result);
  }
  double maxdiff = 0; unsigned dmax = 0;
  for (unsigned i = 0; i < dim; i++) if (diff[i] > maxdiff) { maxdiff = diff[i]; dmax = i; }
  R->split = dmax;
}

/* Gauss-Kronrod 15-point rule, dim 1 (hcubature.c:619-758) */
static void gk15_eval(orule *r, oreg *R)
{
  const unsigned n = 8;
  const double xgk[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000 };
  const double wg[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327 };
  const double wgk[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714 };
  unsigned fdim = r->fdim, np = 0, j;
  double c = R->c[0], hw = R->h[0];
  double *pts = r->pts;
  pts[np++] = c;
  for (j = 0; j < (n-1)/2; j++) { double w = hw*xgk[2*j+1]; pts[np++] = c - w; pts[np++] = c + w; }
  for (j = 0; j < n/2; j++)     { double w = hw*xgk[2*j];   pts[np++] = c - w; pts[np++] = c + w; }
  R->split = 0;
  for (unsigned k = 0; k < np; k++) r->f(1, pts + k, r->fdata, fdim, r->vals + k*fdim);
  for (unsigned k = 0; k < fdim; k++) {
    const double *vk = r->vals + k;
    double rg = vk[0]*wg[n/2 - 1], rk = vk[0]*wgk[n - 1], rabs = fabs(rk), rasc, mean, err;
    unsigned q = 1;
    for (j = 0; j < (n-1)/2; j++) {
      int j2 = 2*j + 1;
      double v = vk[fdim*q] + vk[fdim*q + fdim];
      rg += wg[j]*v; rk += wgk[j2]*v;
      rabs += wgk[j2]*(fabs(vk[fdim*q]) + fabs(vk[fdim*q + fdim]));
      q += 2;
    }
    for (j = 0; j < n/2; j++) {
      int j2 = 2*j;
      rk += wgk[j2]*(vk[fdim*q] + vk[fdim*q + fdim]);
      rabs += wgk[j2]*(fabs(vk[fdim*q]) + fabs(vk[fdim*q + fdim]));
      q += 2;
    }
    R->ee[2*k] = rk*hw;
    mean = rk*0.5;
    rasc = wgk[n - 1]*fabs(vk[0] - mean);
    q = 1;
    for (j = 0; j < (n-1)/2; j++) {
      int j2 = 2*j + 1;
      rasc += wgk[j2]*(fabs(vk[fdim*q] - mean) + fabs(vk[fdim*q + fdim] - mean));
      q += 2;
    }
    for (j = 0; j < n/2; j++) {
      int j2 = 2*j;
      rasc += wgk[j2]*(fabs(vk[fdim*q] - mean) + fabs(vk[fdim*q + fdim] - mean));
      q += 2;
    }
    err = fabs(rk - rg)*hw;
    rabs *= hw; rasc *= hw;
    if (rasc != 0 && err != 0) {
      double scale = pow((200*err/rasc), 1.5);
      err = (scale < 1) ? rasc*scale : rasc;
    }
    if (rabs > DBL_MIN / (50*DBL_EPSILON)) {
      double min_err = 50*DBL_EPSILON*rabs;
      if (min_err > err) err = min_err;
    }
    R->ee[2*k+1] = err;
  }
}

static void reg_eval(orule *r, oreg *R)
{
  if (r->dim == 1) gk15_eval(r, R); else gm_eval(r, R);
  double m = 0;
  for (unsigned k = 0; k < r->fdim; k++) if (R->ee[2*k+1] > m) m = R->ee[2*k+1];
  R->errmax = m;
}

/* binary max-heap of regions keyed on errmax (hcubature.c:772-846) */
typedef struct { size_t n, nalloc; oreg *items; unsigned fdim; double *ee; } oheap;

static void heap_push(oheap *h, oreg hi)
{
  for (unsigned i = 0; i < h->fdim; i++) { h->ee[2*i] += hi.ee[2*i]; h->ee[2*i+1] += hi.ee[2*i+1]; }
  size_t ins = h->n;
  if (++(h->n) > h->nalloc) { h->nalloc = h->n*2; h->items = (oreg *)realloc(h->items, sizeof(oreg)*h->nalloc); }
  while (ins) {
    size_t parent = (ins - 1)/2;
    if (hi.errmax <= h->items[parent].errmax) break;
    h->items[ins] = h->items[parent];
    ins = parent;
  }
  h->items[ins] = hi;
}

static oreg heap_pop(oheap *h)
{
  oreg ret = h->items[0];
  size_t i = 0, n = --(h->n), child;
  h->items[0] = h->items[n];
  while ((child = i*2 + 1) < n) {
    size_t largest;
    if (h->items[child].errmax <= h->items[i].errmax) largest = i; else largest = child;
    if (++child < n && h->items[largest].errmax < h->items[child].errmax) largest = child;
    if (largest == i) break;
    oreg sw = h->items[i]; h->items[i] = h->items[largest]; h->items[i = largest] = sw;
  }
  for (unsigned k = 0; k < h->fdim; k++) { h->ee[2*k] -= ret.ee[2*k]; h->ee[2*k+1] -= ret.ee[2*k+1]; }
  return ret;
}

static int converged_ind(unsigned fdim, const double *ee, double reqAbs, double reqRel)
{
  for (unsigned j = 0; j < fdim; j++)
    if (ee[2*j+1] > reqAbs && ee[2*j+1] > fabs(ee[2*j])*reqRel) return 0;
  return 1;
}

int o_hcubature(unsigned fdim, o_integrand f, void *fdata, unsigned dim,
                const double *xmin, const double *xmax, size_t maxEval,
                double reqAbs, double reqRel, double *val, double *err,
                size_t *numEvalOut)
{
  if (dim < 1 || dim > 3) return -1;
  orule r;
  r.dim = dim; r.fdim = fdim; r.f = f; r.fdata = fdata;
  r.npts = (dim == 1) ? 15 : 1 + 2*2*dim + 2*dim*(dim-1) + (1U << dim);
  r.pts = (double *)malloc(sizeof(double)*r.npts*dim);
  r.vals = (double *)malloc(sizeof(double)*r.npts*fdim);
  oheap h; h.n = 0; h.nalloc = 1; h.fdim = fdim;
  h.items = (oreg *)malloc(sizeof(oreg));
  h.ee = (double *)calloc(2*fdim, sizeof(double));
  /* all region estimates live in a pool; regions are freed at the end */
  size_t poolcap = 64, poolused = 0;
  double **pool = (double **)malloc(sizeof(double *)*poolcap);
#define NEWEE() (poolused == poolcap ? (pool = (double **)realloc(pool, sizeof(double *)*(poolcap *= 2))) : 0, \
                 pool[poolused++] = (double *)malloc(sizeof(double)*2*fdim))
  oreg R0;
  for (unsigned i = 0; i < dim; i++) { R0.c[i] = 0.5*(xmin[i] + xmax[i]); R0.h[i] = 0.5*(xmax[i] - xmin[i]); }
  R0.vol = region_vol(&R0, dim);
  R0.split = 0;
  R0.ee = NEWEE();
  reg_eval(&r, &R0);
  heap_push(&h, R0);
  size_t numEval = r.npts;
  while (numEval < maxEval || !maxEval) {
    if (converged_ind(fdim, h.ee, reqAbs, reqRel)) break;
    oreg A = heap_pop(&h), B;
    /* cut_region (hcubature.c:215-228) */
    unsigned d = A.split;
    B = A;
    A.h[d] *= 0.5;
    A.vol *= 0.5;
    memcpy(B.c, A.c, sizeof(B.c)); memcpy(B.h, A.h, sizeof(B.h));
    B.vol = region_vol(&B, dim);
    A.c[d] -= A.h[d];
    B.c[d] += A.h[d];
    B.ee = NEWEE();
    reg_eval(&r, &A);
    reg_eval(&r, &B);
    heap_push(&h, A);
    heap_push(&h, B);
    numEval += 2*r.npts;
  }
  for (unsigned j = 0; j < fdim; j++) val[j] = err[j] = 0;
  for (size_t i = 0; i < h.n; i++)
    for (unsigned j = 0; j < fdim; j++) { val[j] += h.items[i].ee[2*j]; err[j] += h.items[i].ee[2*j+1]; }
#undef NEWEE
  for (size_t i = 0; i < poolused; i++) free(pool[i]);
  free(pool); free(h.items); free(h.ee); free(r.pts); free(r.vals);
  if (numEvalOut) *numEvalOut = numEval;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Clenshaw-Curtis tables                                              */
/* x: x[0]=1, then for m>=1 at offset 2^(m-1): cos((2i+1)pi/2^(m+1)),  */
/*    i<2^(m-1).                                                        */
/* w: rule of order m (n=2^(m+1) intervals) at offset m+2^m-1, length   */
/*    2^m+1: [w(x=0), w(level-0 node), w(level-1 nodes), ...]          */
/* ------------------------------------------------------------------ */
#define CC_NX (1 << (O_CLENCURT_M))
double o_clencurt_x[CC_NX];
double o_clencurt_w[(O_CLENCURT_M+1) + (1 << (O_CLENCURT_M+1))];
static int cc_ready = 0;

static long double cc_weight(long n, long j)
{
  const long double pi = 3.141592653589793238462643383279502884L;
  long double s = 0;
  for (long k = 1; k <= n/2; k++) {
    long double b = (k == n/2) ? 1.0L : 2.0L;
    s += b / (4.0L*k*k - 1.0L) * cosl(2.0L*pi*(long double)((j*k) % n) / (long double)n);
  }
  long double c = (j == 0 || j == n) ? 1.0L : 2.0L;
  return c / (long double)n * (1.0L - s);
}

void o_clencurt_init(void)
{
  if (cc_ready) return;
  const long double pi = 3.141592653589793238462643383279502884L;
  o_clencurt_x[0] = 1.0;
  for (int m = 1; m <= O_CLENCURT_M - 0 && (1 << (m-1)) < CC_NX; m++)
    for (int i = 0; i < (1 << (m-1)); i++)
      o_clencurt_x[(1 << (m-1)) + i] = (double)cosl((2*i+1)*pi / (long double)(1L << (m+1)));
  /* weights are needed only for the orders the integrators reach; */
  /* fill orders 0..12 (4097-point rule) */
  for (int m = 0; m <= 12; m++) {
    long n = 1L << (m+1);
    double *w = o_clencurt_w + m + (1 << m) - 1;
    w[0] = (double)cc_weight(n, n/2);
    w[1] = (double)cc_weight(n, 0);
    int q = 2;
    for (int l = 1; l <= m; l++)
      for (int i = 0; i < (1 << (l-1)); i++)
        w[q++] = (double)cc_weight(n, (2*i+1)*(n >> (l+1)));
  }
  cc_ready = 1;
}

/* ------------------------------------------------------------------ */
/* p-adaptive cubature (pcubature.c:60-430)                            */
/* ------------------------------------------------------------------ */
#define PMAXDIM 3
typedef struct { unsigned m[PMAXDIM]; unsigned mi; double *val; } ocacheval;
typedef struct { size_t ncache; ocacheval *c; } ovalcache;

static void pc_compute(const unsigned *m, unsigned mi, double *val, size_t *vali,
                       unsigned fdim, o_integrand f, void *fdata, unsigned dim,
                       unsigned id, double *p, const double *xmin, const double *xmax)
{
  if (id == dim) {
    f(dim, p, fdata, fdim, val + *vali);
    *vali += fdim;
    return;
  }
  double c = (xmin[id] + xmax[id])*0.5, r = (xmax[id] - xmin[id])*0.5;
  const double *x = o_clencurt_x + ((id == mi) ? (m[id] ? (1 << (m[id]-1)) : 0) : 0);
  unsigned nx = (id == mi ? (m[id] ? (1u << (m[id]-1)) : 1) : (1u << m[id]));
  if (id != mi) {
    p[id] = c;
    pc_compute(m, mi, val, vali, fdim, f, fdata, dim, id+1, p, xmin, xmax);
  }
  for (unsigned i = 0; i < nx; i++) {
    p[id] = c + r*x[i];
    pc_compute(m, mi, val, vali, fdim, f, fdata, dim, id+1, p, xmin, xmax);
    p[id] = c - r*x[i];
    pc_compute(m, mi, val, vali, fdim, f, fdata, dim, id+1, p, xmin, xmax);
  }
}

static size_t pc_num(const unsigned *m, unsigned mi, unsigned dim)
{
  size_t nval = 1;
  for (unsigned i = 0; i < dim; i++) {
    if (i == mi) nval *= m[i] == 0 ? 2 : (1u << m[i]);
    else nval *= (1u << (m[i]+1)) + 1;
  }
  return nval;
}

static void pc_add(ovalcache *vc, const unsigned *m, unsigned mi, unsigned fdim,
                   o_integrand f, void *fdata, unsigned dim,
                   const double *xmin, const double *xmax)
{
  size_t ic = vc->ncache, vali = 0;
  double p[PMAXDIM];
  vc->c = (ocacheval *)realloc(vc->c, sizeof(ocacheval)*++(vc->ncache));
  vc->c[ic].mi = mi;
  memcpy(vc->c[ic].m, m, sizeof(unsigned)*dim);
  vc->c[ic].val = (double *)malloc(sizeof(double)*fdim*pc_num(m, mi, dim));
  pc_compute(m, mi, vc->c[ic].val, &vali, fdim, f, fdata, dim, 0, p, xmin, xmax);
}

static size_t pc_eval(const unsigned *cm, unsigned cmi, double *cval,
                      const unsigned *m, unsigned md, unsigned fdim, unsigned dim,
                      unsigned id, double weight, double *val)
{
  size_t voff = 0;
  if (id == dim) {
    for (unsigned i = 0; i < fdim; i++) val[i] += cval[i]*weight;
    voff = fdim;
  } else if (m[id] == 0 && id == md) {
    voff = pc_eval(cm, cmi, cval, m, md, fdim, dim, id+1, weight*2, val);
    voff += fdim*(1u << cm[id])*2*pc_num(cm + id+1, cmi - (id+1), dim - (id+1));
  } else {
    unsigned mid = m[id] - (id == md);
    const double *w = o_clencurt_w + mid + (1 << mid) - 1
                      + (id == cmi ? (cm[id] ? 1 + (1 << (cm[id]-1)) : 1) : 0);
    unsigned cnx = (id == cmi ? (cm[id] ? (1u << (cm[id]-1)) : 1) : (1u << cm[id]));
    unsigned nx = cm[id] <= mid ? cnx : (1u << mid);
    if (id != cmi) {
      voff = pc_eval(cm, cmi, cval, m, md, fdim, dim, id+1, weight*w[0], val);
      ++w;
    }
    for (unsigned i = 0; i < nx; i++) {
      voff += pc_eval(cm, cmi, cval + voff, m, md, fdim, dim, id+1, weight*w[i], val);
      voff += pc_eval(cm, cmi, cval + voff, m, md, fdim, dim, id+1, weight*w[i], val);
    }
    voff += (cnx - nx)*fdim*2*pc_num(cm + id+1, cmi - (id+1), dim - (id+1));
  }
  return voff;
}

static void pc_evals(ovalcache *vc, const unsigned *m, unsigned md, unsigned fdim,
                     unsigned dim, double V, double *val)
{
  memset(val, 0, sizeof(double)*fdim);
  for (size_t i = 0; i < vc->ncache; i++)
    if (vc->c[i].mi >= dim || vc->c[i].m[vc->c[i].mi] + (vc->c[i].mi == md) <= m[vc->c[i].mi])
      pc_eval(vc->c[i].m, vc->c[i].mi, vc->c[i].val, m, md, fdim, dim, 0, V, val);
}

int o_pcubature(unsigned fdim, o_integrand f, void *fdata, unsigned dim,
                const double *xmin, const double *xmax, size_t maxEval,
                double reqAbs, double reqRel, double *val, double *err)
{
  o_clencurt_init();
  if (dim == 0) { f(0, xmin, fdata, fdim, val); for (unsigned i = 0; i < fdim; i++) err[i] = 0; return 0; }
  if (dim > PMAXDIM) return -1;
  unsigned m[PMAXDIM] = {0, 0, 0};
  double V = 1;
  size_t numEval = 0;
  int ret = -1;
  ovalcache vc = {0, NULL};
  double *val1 = (double *)malloc(sizeof(double)*fdim);
  for (unsigned i = 0; i < fdim; i++) { val[i] = 0; err[i] = HUGE_VAL; }
  for (unsigned i = 0; i < dim; i++) V *= (xmax[i] - xmin[i])*0.5;
  pc_add(&vc, m, dim, fdim, f, fdata, dim, xmin, xmax);
  for (;;) {
    unsigned mi = 0;
    /* eval_integral (pcubature.c:246-275) */
    pc_evals(&vc, m, dim, fdim, dim, V, val);
    memset(err, 0, sizeof(double)*fdim);
    double maxerr = 0;
    for (unsigned i = 0; i < dim; i++) {
      double emax = 0;
      pc_evals(&vc, m, i, fdim, dim, V, val1);
      for (unsigned j = 0; j < fdim; j++) {
        double e = fabs(val[j] - val1[j]);
        if (e > emax) emax = e;
        if (e > err[j]) err[j] = e;
      }
      if (emax > maxerr) { maxerr = emax; mi = i; }
    }
    int conv = 1;
    for (unsigned j = 0; j < fdim; j++)
      if (err[j] > reqAbs && err[j] > fabs(val[j])*reqRel) { conv = 0; break; }
    if (conv || (numEval > maxEval && maxEval)) { ret = 0; break; }
    m[mi] += 1;
    if (m[mi] > O_CLENCURT_M || m[mi] > 12) break;
    size_t nb = pc_num(m, mi, dim);
    pc_add(&vc, m, mi, fdim, f, fdata, dim, xmin, xmax);
    numEval += nb;
  }
  for (size_t i = 0; i < vc.ncache; i++) free(vc.c[i].val);
  free(vc.c); free(val1);
  return ret;
}
