/*
 * CPU oracle for the MLFMAxwell hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in plain single-threaded C, the reference's algorithm for:
 *   - RWG construction           modRWG.f90:8-174, 270-328
 *   - MLFMA setup                modMLFMA.f90:113-275, 453-1391, 1594-1634
 *   - MLFMA matvec               modMLFMA.f90:1783-1946
 *   - near-field MoM entries     modMoM.f90:89-135, modGlobalMetod.f90:472-487
 *   - transfer functions         modMLFMA.f90:964-994 with hankel/sphj/sphy
 *                                (modGlobalMetod.f90:493-979) and legendre
 *                                (modGlobalMetod.f90:988-1016)
 *   - ILU0 / triangular solves   modIterativo.f90:1805-1976, 504-560
 *   - CGS                        modIterativo.f90:331-493
 *   - plane wave                 main.f90:1488-1525
 *   - RCS                        modRCS.f90:23-153
 *
 * Floating-point details that decide bitwise agreement with a flang build of
 * the reference are reproduced on purpose:
 *   - single-precision literals promoted to double: (1./3.), (2./3.),
 *     (1./72.), sqrt(3.), (1./3.) in the multipole count, log10(2.);
 *   - real**integer is binary exponentiation (flang's lowering);
 *   - NORM2 is flang's one-pass scaled algorithm;
 *   - DOT_PRODUCT of 3-vectors sums left to right;
 *   - near-field values and the ILU0 factor are stored as complex(4).
 * The binomial Legendre form (modGlobalMetod.f90:988) loses ~1e-6 absolute at
 * degree 38 to cancellation, so its inputs must be bit-identical to the
 * reference's for transfer functions to agree beyond that level.
 *
 * This file is never linked into the product library.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;
typedef float complex cplxf;

/* modGlobalParam.f90:12-17 */
static const double MU = 1.2566370614E-6;
static const double EPSI = 8.8541878176E-12;
static const double PI = 3.141592653589793;
static double g_lambda = 1.0, g_w, g_kappa, g_vp;

void or_set_lambda(double lambda) {
  /* setlambda, modGlobalParam.f90:53-61 */
  g_vp = 1.0 / sqrt(MU * EPSI);
  g_lambda = lambda;
  g_w = 2.0 * PI * g_vp / lambda;
  g_kappa = 2.0 * PI / lambda;
}
double or_get_w(void) { return g_w; }

/* ---------------------------------------------------------------- math */

/* flang lowers real(8)**integer to binary exponentiation. */
static double ipow(double x, int n) {
  double r = 1.0, b = x;
  unsigned m = (unsigned)n;
  while (m) {
    if (m & 1u) r *= b;
    b *= b;
    m >>= 1;
  }
  return r;
}

/* flang runtime NORM2: one pass, scaled by the running maximum. */
double or_norm2(const double *v) {
  double m = 0.0, s = 0.0;
  for (int i = 0; i < 3; ++i) {
    double a = fabs(v[i]);
    if (m == 0.0) {
      m = a;
    } else if (a > m) {
      double t = m / a, tsq = t * t;
      s *= tsq;
      s += tsq;
      m = a;
    } else {
      double t = a / m;
      s += t * t;
    }
  }
  return m * sqrt(1.0 + s);
}

static double dot3(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* factorial, modGlobalMetod.f90:1004-1016 */
static double factorial(int x) {
  double r = 1.0;
  for (int i = 1; i <= x; ++i) r = r * (double)i;
  return r;
}

/* legendre, modGlobalMetod.f90:988-1002 (binomial form, kept as is). */
double or_legendre(int n, double x) {
  double res = 0.0;
  for (int i = 0; i <= n; ++i) {
    double c = factorial(n) / (factorial(i) * factorial(n - i));
    res = res + ((c * c) * ipow(x + 1.0, n - i)) * ipow(x - 1.0, i);
  }
  /* 0.5**n in real(4) is an exact power of two */
  return ldexp(1.0, -n) * res;
}

/* envj / msta1 / msta2, modGlobalMetod.f90:548-772 (Zhang & Jin). */
static double envj(int n, double x) {
  return 0.5 * log10(6.28 * n) - n * log10(1.36 * x / n);
}
static int msta1(double x, int mp) {
  double a0 = fabs(x);
  int n0 = (int)(1.1 * a0) + 1;
  double f0 = envj(n0, a0) - mp;
  int n1 = n0 + 5;
  double f1 = envj(n1, a0) - mp;
  int nn = n1;
  for (int it = 1; it <= 20; ++it) {
    nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
    double f = envj(nn, a0) - mp;
    if (abs(nn - n1) < 1) break;
    n0 = n1;
    f0 = f1;
    n1 = nn;
    f1 = f;
  }
  return nn;
}
static int msta2(double x, int n, int mp) {
  double a0 = fabs(x);
  double hmp = 0.5 * mp;
  double ejn = envj(n, a0);
  double obj;
  int n0;
  if (ejn <= hmp) {
    obj = mp;
    n0 = (int)(1.1 * a0);
  } else {
    obj = hmp + ejn;
    n0 = n;
  }
  double f0 = envj(n0, a0) - obj;
  int n1 = n0 + 5;
  double f1 = envj(n1, a0) - obj;
  int nn = n1;
  for (int it = 1; it <= 20; ++it) {
    nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
    double f = envj(nn, a0) - obj;
    if (abs(nn - n1) < 1) break;
    n0 = n1;
    f0 = f1;
    n1 = nn;
    f1 = f;
  }
  return nn + 10;
}
/* sphj, modGlobalMetod.f90:865-979: returns j_n(x) as the routine computes
 * it when called with order n (sphBesselJ). */
static double sphj_n(int n, double x) {
  if (fabs(x) <= 1.0e-100) return n == 0 ? 1.0 : 0.0;
  double sj0 = sin(x) / x;
  double sj1 = (sj0 - cos(x)) / x;
  if (n < 2) return n == 0 ? sj0 : sj1;
  double sa = sj0, sb = sj1;
  int nm = n;
  int m = msta1(x, 200);
  if (m < n)
    nm = m;
  else
    m = msta2(x, n, 15);
  double *sj = (double *)calloc((size_t)n + 1, sizeof(double));
  double f0 = 0.0, f1 = 1.0 - 100.0; /* the reference writes 1.0D+00-100 */
  double f = 0.0;
  for (int k = m; k >= 0; --k) {
    f = (2.0 * k + 3.0) * f1 / x - f0;
    if (k <= nm) sj[k] = f;
    f0 = f1;
    f1 = f;
  }
  double cs = (fabs(sa) <= fabs(sb)) ? sb / f0 : sa / f;
  for (int k = 0; k <= nm; ++k) sj[k] = cs * sj[k];
  double r = sj[n];
  free(sj);
  return r;
}
/* sphy, modGlobalMetod.f90:777-859 */
static double sphy_n(int n, double x) {
  if (x < 1.0e-60) return -1.0e300;
  double sy0 = -cos(x) / x;
  double sy1 = (sy0 - sin(x)) / x;
  if (n == 0) return sy0;
  if (n == 1) return sy1;
  double f0 = sy0, f1 = sy1, f = 0.0;
  for (int k = 2; k <= n; ++k) {
    f = (2.0 * k - 1.0) * f1 / x - f0;
    if (1.0e300 <= fabs(f)) return f;
    f0 = f1;
    f1 = f;
  }
  return f;
}
void or_hankel(int n, double x, double *h) {
  h[0] = sphj_n(n, x);
  h[1] = -sphy_n(n, x);
}

/* green, modGlobalMetod.f90:472-487 */
static cplx green(const double *pm, const double *pn) {
  double d[3] = {pm[0] - pn[0], pm[1] - pn[1], pm[2] - pn[2]};
  double r = sqrt(dot3(d, d));
  double ph = -((g_w * sqrt(MU * EPSI)) * r);
  return (cos(ph) + I * sin(ph)) / r;
}

/* ---------------------------------------------------------------- RWG */

int or_rwg_build(int np, const double *p, int nt, const int *t, int *e_p,
                 int *e_t, int *e_po, double *e_long, double *e_centro) {
  (void)np;
  int nall = 3 * nt;
  int *lt = (int *)malloc(sizeof(int) * 2 * (size_t)nall);
  /* llenar_lados, modRWG.f90:8-29 */
  for (int it = 0; it < nt; ++it) {
    lt[2 * (3 * it + 0) + 0] = t[3 * it + 0];
    lt[2 * (3 * it + 0) + 1] = t[3 * it + 1];
    lt[2 * (3 * it + 1) + 0] = t[3 * it + 0];
    lt[2 * (3 * it + 1) + 1] = t[3 * it + 2];
    lt[2 * (3 * it + 2) + 0] = t[3 * it + 1];
    lt[2 * (3 * it + 2) + 1] = t[3 * it + 2];
  }
  /* Buscar_L_Comun returns the first later half-edge with the same vertex
   * pair (either orientation).  To avoid the O(E^2) scan, half-edges are
   * sorted by (min,max,index) once; the next equal key in that order is
   * exactly the first later match. */
  long long *key = (long long *)malloc(sizeof(long long) * (size_t)nall);
  int *ord = (int *)malloc(sizeof(int) * (size_t)nall);
  int *next = (int *)malloc(sizeof(int) * (size_t)nall);
  for (int i = 0; i < nall; ++i) {
    int a = lt[2 * i], b = lt[2 * i + 1];
    int lo = a < b ? a : b, hi = a < b ? b : a;
    key[i] = ((long long)lo << 32) | (unsigned)hi;
    ord[i] = i;
  }
  /* simple merge sort of ord by (key, index) */
  int *tmp = (int *)malloc(sizeof(int) * (size_t)nall);
  for (int w = 1; w < nall; w *= 2) {
    for (int lo = 0; lo < nall; lo += 2 * w) {
      int mid = lo + w < nall ? lo + w : nall;
      int hi = lo + 2 * w < nall ? lo + 2 * w : nall;
      int a = lo, b = mid, o = lo;
      while (a < mid && b < hi) {
        int ia = ord[a], ib = ord[b];
        int take_a = key[ia] < key[ib] || (key[ia] == key[ib] && ia < ib);
        tmp[o++] = take_a ? ord[a++] : ord[b++];
      }
      while (a < mid) tmp[o++] = ord[a++];
      while (b < hi) tmp[o++] = ord[b++];
    }
    memcpy(ord, tmp, sizeof(int) * (size_t)nall);
  }
  for (int i = 0; i < nall; ++i) next[i] = -1;
  for (int s = 0; s + 1 < nall; ++s)
    if (key[ord[s]] == key[ord[s + 1]]) next[ord[s]] = ord[s + 1];
  free(tmp);
  free(ord);
  free(key);

  /* llenar_todo, modRWG.f90:59-131 */
  char *dead = (char *)calloc((size_t)nall, 1);
  int ne = 0;
  for (int i = 0; i < nall; ++i) {
    if (dead[i]) continue;
    /* first later live match: walk the equal-key chain skipping dead ones */
    int bq = next[i];
    while (bq >= 0 && dead[bq]) bq = next[bq];
    if (bq >= 0) {
      e_p[2 * ne] = lt[2 * i];
      e_p[2 * ne + 1] = lt[2 * i + 1];
      dead[bq] = 1;
      int x = i / 3 + 1, y = bq / 3 + 1; /* 1-based triangles */
      e_t[2 * ne] = x;
      e_t[2 * ne + 1] = y;
      const int *tx = t + 3 * (x - 1), *ty = t + 3 * (y - 1);
      int a = e_p[2 * ne], b = e_p[2 * ne + 1];
      e_po[2 * ne] = (tx[0] != a && tx[0] != b) ? tx[0]
                     : (tx[1] != a && tx[1] != b) ? tx[1]
                                                  : tx[2];
      e_po[2 * ne + 1] = (ty[0] != a && ty[0] != b) ? ty[0]
                         : (ty[1] != a && ty[1] != b) ? ty[1]
                                                      : ty[2];
      ++ne;
    } else {
      dead[i] = 1;
    }
  }
  free(dead);
  free(next);
  free(lt);
  /* parametros_rwg, modRWG.f90:162-166 */
  for (int i = 0; i < ne; ++i) {
    const double *a = p + 3 * (e_p[2 * i] - 1), *b = p + 3 * (e_p[2 * i + 1] - 1);
    double v[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    e_long[i] = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) e_centro[3 * i + k] = 0.5 * (a[k] + b[k]);
  }
  return ne;
}

void or_tri_geometry(int np, const double *p, int nt, const int *t,
                     double *t_area, double *t_baric, double *t_baric_sub) {
  (void)np;
  const double third = (double)(1.0f / 3.0f), two3 = (double)(2.0f / 3.0f);
  for (int i = 0; i < nt; ++i) {
    const double *p1 = p + 3 * (t[3 * i] - 1), *p2 = p + 3 * (t[3 * i + 1] - 1),
                 *p3 = p + 3 * (t[3 * i + 2] - 1);
    double v[11][3];
    for (int k = 0; k < 3; ++k) {
      t_baric[3 * i + k] = third * ((p1[k] + p2[k]) + p3[k]);
      v[1][k] = p1[k];
      v[2][k] = p2[k];
      v[3][k] = p3[k];
      v[4][k] = p1[k] + third * (p2[k] - p1[k]);
      v[5][k] = p1[k] + two3 * (p2[k] - p1[k]);
      v[6][k] = p2[k] + third * (p3[k] - p2[k]);
      v[7][k] = p2[k] + two3 * (p3[k] - p2[k]);
      v[8][k] = p3[k] + third * (p1[k] - p3[k]);
      v[9][k] = p3[k] + two3 * (p1[k] - p3[k]);
      v[10][k] = t_baric[3 * i + k];
    }
    static const int sub[9][3] = {{1, 4, 9}, {4, 5, 10}, {5, 2, 6},
                                  {6, 10, 7}, {7, 8, 3}, {7, 8, 10},
                                  {8, 9, 10}, {9, 4, 10}, {5, 6, 10}};
    for (int s = 0; s < 9; ++s)
      for (int k = 0; k < 3; ++k)
        t_baric_sub[27 * i + 3 * s + k] =
            third * ((v[sub[s][0]][k] + v[sub[s][1]][k]) + v[sub[s][2]][k]);
    double ba[3], ca[3], cr[3];
    for (int k = 0; k < 3; ++k) {
      ba[k] = p1[k] - p2[k];
      ca[k] = p1[k] - p3[k];
    }
    cr[0] = ba[1] * ca[2] - ba[2] * ca[1];
    cr[1] = ba[2] * ca[0] - ba[0] * ca[2];
    cr[2] = ba[0] * ca[1] - ba[1] * ca[0];
    t_area[i] = 0.5 * sqrt((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
  }
}

/* ---------------------------------------------------------------- octree */

typedef struct {
  double c[3];
  int padre; /* 1-based index in the parent level */
  int np, *prop;
  int ncer, *cer;
  int nfar, *far_box, *far_tr; /* 1-based */
} Box;

typedef struct {
  Box *b;
  int nb, cap, npadres;
  double lado;
  int nmult, nth, nph, nk;
  double *kap, *ath, *aph, *ang, *pes;
  int nT;
  cplx *T;       /* (nk, nT) */
  double *rab;   /* (3, nT) first-seen box offset of each transfer function */
  double *wint;  /* (9*nk) for the interpolation to this level from level+1 */
  int *iint;     /* 1-based fine-sample indices */
  cplx *C, *B, *A;
} Level;

struct or_mlfma {
  int L;
  Level *lv; /* lv[1..L] */
  int ne, nt, np, fp64_near;
  const double *e_centro, *e_long, *t_area, *t_baric, *t_baric_sub, *p_coord,
      *t_normal;
  const int *e_t, *e_po, *t_p, *e_p;
  cplx *rad; /* funcRAD (nk_L, 2, ne) */
  long long nnz;
  int *ia, *ja;
  cplxf *z4;
  cplx *z8;
  int *leaf; /* leaf box (1-based) of each basis */
};

static Box *add_box(Level *lv) {
  if (lv->nb == lv->cap) {
    lv->cap = lv->cap ? 2 * lv->cap : 64;
    lv->b = (Box *)realloc(lv->b, sizeof(Box) * (size_t)lv->cap);
  }
  Box *b = &lv->b[lv->nb++];
  memset(b, 0, sizeof(Box));
  return b;
}

/* calcRad, non-maxPrecision branch (modMLFMA.f90:669-699). */
static void calc_rad(or_mlfma *h, const double *c, const int *fb, int nb) {
  Level *lv = &h->lv[h->L];
  double ka = (2.0 * PI) / g_lambda;
  for (int i = 0; i < nb; ++i) {
    int f = fb[i] - 1;
    const double *bp = h->t_baric + 3 * (h->e_t[2 * f] - 1);
    const double *vp = h->p_coord + 3 * (h->e_po[2 * f] - 1);
    const double *bm = h->t_baric + 3 * (h->e_t[2 * f + 1] - 1);
    const double *vm = h->p_coord + 3 * (h->e_po[2 * f + 1] - 1);
    double lon = h->e_long[f];
    double rp[3], rm[3], dp_[3], dm_[3];
    for (int k = 0; k < 3; ++k) {
      rp[k] = bp[k] - vp[k];
      rm[k] = vm[k] - bm[k];
      dp_[k] = bp[k] - c[k];
      dm_[k] = bm[k] - c[k];
    }
    for (int j = 0; j < lv->nk; ++j) {
      const double *vk = lv->kap + 3 * j;
      cplx ep = cexp(I * (ka * dot3(vk, dp_)));
      cplx em = cexp(I * (ka * dot3(vk, dm_)));
      cplx vt[3];
      for (int k = 0; k < 3; ++k) vt[k] = (ep * rp[k] + em * rm[k]) * lon * 0.5;
      const double *at = lv->ath + 3 * j, *ap = lv->aph + 3 * j;
      h->rad[((size_t)f * 2 + 0) * lv->nk + j] = (vt[0] * at[0] + vt[1] * at[1]) + vt[2] * at[2];
      h->rad[((size_t)f * 2 + 1) * lv->nk + j] = (vt[0] * ap[0] + vt[1] * ap[1]) + vt[2] * ap[2];
    }
  }
}

/* geomMLFMA / dividir_en_8_hijos / buscarFuncBases (modMLFMA.f90:497-833). */
static void geom(or_mlfma *h, const double *c, double lcubo, const int *vb,
                 int nb, int nivel) {
  if (nivel == h->L) {
    calc_rad(h, c, vb, nb);
    return;
  }
  static const int sx[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  static const int sy[8] = {-1, 1, -1, 1, -1, 1, -1, 1};
  static const int sz[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  double lh = lcubo / 2.0, lq = lcubo / 4.0;
  int *vbn = (int *)malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  memcpy(vbn, vb, sizeof(int) * (size_t)nb);
  int nbn = nb;
  Level *nx = &h->lv[nivel + 1];
  nx->npadres += 1;
  int mypadre = nx->npadres;
  for (int i = 0; i < 8; ++i) {
    double ch[3] = {sx[i] < 0 ? c[0] - lq : c[0] + lq,
                    sy[i] < 0 ? c[1] - lq : c[1] + lq,
                    sz[i] < 0 ? c[2] - lq : c[2] + lq};
    double q = lh / 2.0;
    double lim[6] = {ch[0] + q, ch[0] - q, ch[1] + q, ch[1] - q, ch[2] + q, ch[2] - q};
    int *enc = (int *)malloc(sizeof(int) * (size_t)(nbn > 0 ? nbn : 1));
    int *rest = (int *)malloc(sizeof(int) * (size_t)(nbn > 0 ? nbn : 1));
    int ne_ = 0, nr = 0;
    for (int k = 0; k < nbn; ++k) {
      const double *e = h->e_centro + 3 * (vbn[k] - 1);
      if (e[0] <= lim[0] && e[0] >= lim[1] && e[1] <= lim[2] && e[1] >= lim[3] &&
          e[2] <= lim[4] && e[2] >= lim[5])
        enc[ne_++] = vbn[k];
      else
        rest[nr++] = vbn[k];
    }
    if (ne_ > 0) {
      Box *b = add_box(nx);
      memcpy(b->c, ch, sizeof(ch));
      b->padre = mypadre;
      b->np = ne_;
      b->prop = (int *)malloc(sizeof(int) * (size_t)ne_);
      memcpy(b->prop, enc, sizeof(int) * (size_t)ne_);
      geom(h, ch, lh, enc, ne_, nivel + 1);
      memcpy(vbn, rest, sizeof(int) * (size_t)nr);
      nbn = nr;
    }
    free(enc);
    free(rest);
  }
  free(vbn);
}

/* llenar_kappas (modMLFMA.f90:453-495) */
static void fill_kappas(Level *lv) {
  int nk = lv->nk;
  lv->kap = (double *)malloc(sizeof(double) * 3 * (size_t)nk);
  lv->ath = (double *)malloc(sizeof(double) * 3 * (size_t)nk);
  lv->aph = (double *)malloc(sizeof(double) * 3 * (size_t)nk);
  lv->ang = (double *)malloc(sizeof(double) * 2 * (size_t)nk);
  lv->pes = (double *)malloc(sizeof(double) * (size_t)nk);
  double dphi = 2.0 * PI / lv->nph, dth = PI / lv->nth;
  int cont = 0;
  for (int j = 1; j <= lv->nph; ++j) {
    double phi = (dphi / 2.0) * (double)((2 * j) - 1);
    for (int k = 1; k <= lv->nth; ++k) {
      double th = (dth / 2.0) * (double)((2 * k) - 1);
      double *kp = lv->kap + 3 * cont, *at = lv->ath + 3 * cont, *ap = lv->aph + 3 * cont;
      kp[0] = sin(th) * cos(phi);
      kp[1] = sin(th) * sin(phi);
      kp[2] = cos(th);
      at[0] = cos(th) * cos(phi);
      at[1] = cos(th) * sin(phi);
      at[2] = -sin(th);
      ap[0] = -sin(phi);
      ap[1] = cos(phi);
      ap[2] = 0.0;
      lv->ang[2 * cont] = th;
      lv->ang[2 * cont + 1] = phi;
      lv->pes[cont] = (cos((k - 1) * dth) - cos(k * dth)) * dphi;
      ++cont;
    }
  }
}

/* calc_functransf (modMLFMA.f90:964-994) */
static void calc_functransf(const Level *lv, const double *rab, cplx *v) {
  int Ltrue = lv->nmult;
  double modul = or_norm2(rab);
  double kD = (2.0 * PI / g_lambda) * modul;
  int kDint = (int)kD;
  if (kDint < Ltrue) Ltrue = kDint;
  double u[3] = {rab[0] / modul, rab[1] / modul, rab[2] / modul};
  double *hr = (double *)malloc(sizeof(double) * (size_t)(Ltrue + 1));
  double *hi = (double *)malloc(sizeof(double) * (size_t)(Ltrue + 1));
  for (int j = 0; j <= Ltrue; ++j) {
    double hh[2];
    or_hankel(j, kD, hh);
    hr[j] = hh[0];
    hi[j] = hh[1];
  }
  for (int i = 0; i < lv->nk; ++i) {
    double x = dot3(lv->kap + 3 * i, u);
    double vr = 0.0, vi = 0.0;
    for (int j = 0; j <= Ltrue; ++j) {
      double r = (double)(2.0f * (float)j + 1.0f);
      double P = or_legendre(j, x);
      double br, bi;
      switch ((j + 1) & 3) { /* (-i)**(j+1) * (2j+1) * h_j */
        case 0: br = r * hr[j]; bi = r * hi[j]; break;
        case 1: br = r * hi[j]; bi = -(r * hr[j]); break;
        case 2: br = -(r * hr[j]); bi = -(r * hi[j]); break;
        default: br = -(r * hi[j]); bi = r * hr[j]; break;
      }
      vr = vr + br * P;
      vi = vi + bi * P;
    }
    double kk = 2.0 * PI / g_lambda;
    v[i] = ((vr * kk) + I * (vi * kk)) / (4.0 * PI);
  }
  free(hr);
  free(hi);
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* llenar_functransf (modMLFMA.f90:835-960) */
static void fill_transfers(or_mlfma *h) {
  double tol = g_lambda / 100.0;
  int L = h->L;
  for (int n = 3; n <= L; ++n) {
    Level *lv = &h->lv[n], *pl = &h->lv[n - 1];
    double rabT[316][3];
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)lv->nk * 316);
    int cont = 1;
    int *near_buf = (n == L) ? (int *)malloc(sizeof(int) * (size_t)h->ne) : NULL;
    int auxlej[2][316];
    for (int i = 0; i < lv->nb; ++i) {
      Box *bi = &lv->b[i];
      int nlej = 0, ncer = 0;
      for (int j = 0; j < lv->nb; ++j) {
        Box *bj = &lv->b[j];
        const double *ci = pl->b[bi->padre - 1].c, *cj = pl->b[bj->padre - 1].c;
        double rp[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]};
        if (or_norm2(rp) < (2 * pl->lado - tol)) {
          double rh[3] = {bi->c[0] - bj->c[0], bi->c[1] - bj->c[1], bi->c[2] - bj->c[2]};
          if (or_norm2(rh) >= (2 * lv->lado - tol)) {
            nlej++;
            for (int k = 1; k <= cont; ++k) {
              double d0 = (k < cont ? rabT[k - 1][0] : 0.0) - rh[0];
              double d1 = (k < cont ? rabT[k - 1][1] : 0.0) - rh[1];
              double d2 = (k < cont ? rabT[k - 1][2] : 0.0) - rh[2];
              if (fabs(d0) < tol && fabs(d1) < tol && fabs(d2) < tol) {
                auxlej[0][nlej - 1] = j + 1;
                auxlej[1][nlej - 1] = k;
                break;
              } else if (k == cont) {
                if (cont > 316) {
                  fprintf(stderr, "oracle: more than 316 transfer vectors\n");
                  abort();
                }
                memcpy(rabT[cont - 1], rh, sizeof(rh));
                calc_functransf(lv, rh, T + (size_t)(cont - 1) * lv->nk);
                auxlej[0][nlej - 1] = j + 1;
                auxlej[1][nlej - 1] = cont;
                cont++;
                break;
              }
            }
          } else if (n == L) {
            memcpy(near_buf + ncer, bj->prop, sizeof(int) * (size_t)bj->np);
            ncer += bj->np;
          }
        }
      }
      bi->nfar = nlej;
      bi->far_box = (int *)malloc(sizeof(int) * (size_t)(nlej ? nlej : 1));
      bi->far_tr = (int *)malloc(sizeof(int) * (size_t)(nlej ? nlej : 1));
      for (int k = 0; k < nlej; ++k) {
        bi->far_box[k] = auxlej[0][k];
        bi->far_tr[k] = auxlej[1][k];
      }
      if (n == L) {
        qsort(near_buf, (size_t)ncer, sizeof(int), cmp_int);
        bi->ncer = ncer;
        bi->cer = (int *)malloc(sizeof(int) * (size_t)(ncer ? ncer : 1));
        memcpy(bi->cer, near_buf, sizeof(int) * (size_t)ncer);
        h->nnz += (long long)bi->np * ncer;
      }
    }
    lv->nT = cont - 1;
    lv->T = (cplx *)realloc(T, sizeof(cplx) * (size_t)lv->nk * (size_t)(lv->nT ? lv->nT : 1));
    lv->rab = (double *)malloc(sizeof(double) * 3 * (size_t)(lv->nT ? lv->nT : 1));
    memcpy(lv->rab, rabT, sizeof(double) * 3 * (size_t)lv->nT);
    free(near_buf);
  }
}

/* hallarKcercanos (modMLFMA.f90:1096-1315) */
static void hallar_kcercanos(const Level *lc, const Level *lf, int indice, int V[9]) {
  long long nph = lf->nph, nth = lf->nth;
  double phi = lc->ang[2 * (indice - 1) + 1], theta = lc->ang[2 * (indice - 1)];
#define PHI(p) (lf->ang[2 * (((p)-1) * nth) + 1])
#define THE(p) (lf->ang[2 * ((p)-1)])
  long long techo = 1, piso = nph, posmid;
  double del = PHI(2) - PHI(1);
  for (;;) {
    posmid = (techo + piso) / 2;
    double valorpos = PHI(posmid);
    if (posmid == techo || posmid == piso) {
      posmid = (fabs(phi - PHI(techo)) < fabs(phi - PHI(piso))) ? techo : piso;
      break;
    }
    if (valorpos == phi) break;
    if (del > 0) {
      if (valorpos > phi) piso = posmid;
      else if (valorpos < phi) techo = posmid;
    } else {
      if (valorpos > phi) techo = posmid;
      else if (valorpos < phi) piso = posmid;
    }
  }
  long long pP = posmid, pPmay, pPmen;
  if (del > 0) {
    if (pP == 1) { pPmay = pP + 1; pPmen = nph; }
    else if (pP == nph) { pPmay = 1; pPmen = pP - 1; }
    else { pPmen = pP - 1; pPmay = pP + 1; }
  } else {
    if (pP == 1) { pPmen = pP + 1; pPmay = nph; }
    else if (pP == nph) { pPmen = 1; pPmay = pP - 1; }
    else { pPmay = pP - 1; pPmen = pP + 1; }
  }
  techo = 1;
  piso = nth;
  del = THE(2) - THE(1);
  for (;;) {
    posmid = (techo + piso) / 2;
    double valorpos = THE(posmid);
    if (posmid == techo || posmid == piso) {
      posmid = (fabs(theta - THE(techo)) < fabs(theta - THE(piso))) ? techo : piso;
      break;
    }
    if (valorpos == theta) break;
    if (del > 0) {
      if (valorpos > theta) piso = posmid;
      else if (valorpos < theta) techo = posmid;
    } else {
      if (valorpos > theta) techo = posmid;
      else if (valorpos < theta) piso = posmid;
    }
  }
  long long pT = posmid, pTmay, pTmen;
  if (del > 0) {
    if (pT == 1) { pTmen = nth; pTmay = pT + 1; }
    else if (pT == nth) { pTmen = pT - 1; pTmay = 1; }
    else { pTmay = pT + 1; pTmen = pT - 1; }
  } else {
    if (pT == 1) { pTmay = nth; pTmen = pT + 1; }
    else if (pT == nth) { pTmay = pT - 1; pTmen = 1; }
    else { pTmen = pT + 1; pTmay = pT - 1; }
  }
#undef PHI
#undef THE
  V[0] = (int)(nth * (pP - 1) + pT);
  V[1] = (int)(nth * (pP - 1) + pTmen);
  V[2] = (int)(nth * (pP - 1) + pTmay);
  V[3] = (int)(nth * (pPmen - 1) + pT);
  V[4] = (int)(nth * (pPmen - 1) + pTmen);
  V[5] = (int)(nth * (pPmen - 1) + pTmay);
  V[6] = (int)(nth * (pPmay - 1) + pT);
  V[7] = (int)(nth * (pPmay - 1) + pTmen);
  V[8] = (int)(nth * (pPmay - 1) + pTmay);
}

/* Wlagrange3 (modMLFMA.f90:1318-1341) */
static void wlagrange3(double x, const double *xm, double *wi) {
  for (int i = 0; i < 3; ++i) {
    double ww = 1.0;
    for (int j = 0; j < 3; ++j)
      if (j != i) ww = ww * ((x - xm[j]) / (xm[i] - xm[j]));
    wi[i] = ww;
  }
}

/* llenarMinterp (modMLFMA.f90:998-1092) */
static void fill_interp(or_mlfma *h) {
  for (int n = 3; n <= h->L - 1; ++n) {
    Level *lc = &h->lv[n], *lf = &h->lv[n + 1];
    lc->wint = (double *)malloc(sizeof(double) * 9 * (size_t)lc->nk);
    lc->iint = (int *)malloc(sizeof(int) * 9 * (size_t)lc->nk);
    int cont = 0;
    for (int i = 1; i <= lc->nk; ++i) {
      double xt = lc->ang[2 * (i - 1)], xp = lc->ang[2 * (i - 1) + 1];
      int V[9];
      hallar_kcercanos(lc, lf, i, V);
      double mt[3] = {lf->ang[2 * (V[0] - 1)], lf->ang[2 * (V[1] - 1)], lf->ang[2 * (V[2] - 1)]};
      double mp[3] = {lf->ang[2 * (V[0] - 1) + 1], lf->ang[2 * (V[3] - 1) + 1],
                      lf->ang[2 * (V[6] - 1) + 1]};
      double wt[3], wp[3];
      wlagrange3(xt, mt, wt);
      wlagrange3(xp, mp, wp);
      for (int m1 = 0; m1 < 3; ++m1)
        for (int m2 = 0; m2 < 3; ++m2) lc->wint[cont++] = wt[m2] * wp[m1];
      for (int k = 0; k < 9; ++k) lc->iint[cont - 9 + k] = V[k];
    }
  }
}

/* Zmn_MoM (modMoM.f90:89-135) */
static cplx zmn(const or_mlfma *h, int m1, int n1) {
  int m = m1 - 1, n = n1 - 1;
  cplx AP[3] = {0, 0, 0}, AM[3] = {0, 0, 0}, PhiP = 0, PhiM = 0;
  const double *tm1 = h->t_baric + 3 * (h->e_t[2 * m] - 1);
  const double *tm2 = h->t_baric + 3 * (h->e_t[2 * m + 1] - 1);
  const double *sn1 = h->t_baric_sub + 27 * (h->e_t[2 * n] - 1);
  const double *sn2 = h->t_baric_sub + 27 * (h->e_t[2 * n + 1] - 1);
  const double *pn1 = h->p_coord + 3 * (h->e_po[2 * n] - 1);
  const double *pn2 = h->p_coord + 3 * (h->e_po[2 * n + 1] - 1);
  for (int i = 0; i < 9; ++i) {
    cplx g11 = green(tm1, sn1 + 3 * i), g12 = green(tm1, sn2 + 3 * i);
    cplx g21 = green(tm2, sn1 + 3 * i), g22 = green(tm2, sn2 + 3 * i);
    double rho1[3], rho2[3];
    for (int k = 0; k < 3; ++k) {
      rho1[k] = sn1[3 * i + k] - pn1[k];
      rho2[k] = pn2[k] - sn2[3 * i + k];
    }
    for (int k = 0; k < 3; ++k) {
      AP[k] = AP[k] + (g11 * rho1[k]) + (g12 * rho2[k]);
      AM[k] = AM[k] + (g21 * rho1[k]) + (g22 * rho2[k]);
    }
    PhiP = PhiP + (g11 - g12);
    PhiM = PhiM + (g21 - g22);
  }
  double c72 = (double)(1.0f / 72.0f);
  double ln = h->e_long[n], lm = h->e_long[m];
  for (int k = 0; k < 3; ++k) {
    AP[k] = AP[k] * (MU / PI) * c72 * ln;
    AM[k] = AM[k] * (MU / PI) * c72 * ln;
  }
  cplx den = I * 36.0 * PI * g_w * EPSI;
  PhiP = (-1.0) * PhiP * ln / den;
  PhiM = (-1.0) * PhiM * ln / den;
  const double *pm1 = h->p_coord + 3 * (h->e_po[2 * m] - 1);
  const double *pm2 = h->p_coord + 3 * (h->e_po[2 * m + 1] - 1);
  cplx s1 = 0, s2 = 0;
  for (int k = 0; k < 3; ++k) {
    s1 = s1 + AP[k] * (tm1[k] - pm1[k]);
    s2 = s2 + AM[k] * (-tm2[k] + pm2[k]);
  }
  cplx PintA = (I * 0.5 * g_w * lm) * (s1 + s2);
  cplx PintPhi = lm * (PhiM - PhiP);
  return PintPhi + PintA;
}

void or_zmn(or_mlfma *h, int m, int n, double *z) {
  cplx v = zmn(h, m, n);
  z[0] = creal(v);
  z[1] = cimag(v);
}

/* llenarZesparcida (modMLFMA.f90:1594-1634) */
static void fill_near(or_mlfma *h) {
  Level *lv = &h->lv[h->L];
  h->leaf = (int *)malloc(sizeof(int) * (size_t)h->ne);
  for (int c = 0; c < lv->nb; ++c)
    for (int k = 0; k < lv->b[c].np; ++k) h->leaf[lv->b[c].prop[k] - 1] = c + 1;
  h->ia = (int *)malloc(sizeof(int) * (size_t)(h->ne + 1));
  h->ja = (int *)malloc(sizeof(int) * (size_t)h->nnz);
  if (h->fp64_near)
    h->z8 = (cplx *)malloc(sizeof(cplx) * (size_t)h->nnz);
  else
    h->z4 = (cplxf *)malloc(sizeof(cplxf) * (size_t)h->nnz);
  long long cont = 0;
  for (int m = 1; m <= h->ne; ++m) {
    Box *b = &lv->b[h->leaf[m - 1] - 1];
    h->ia[m - 1] = (int)(cont + 1);
    for (int j = 0; j < b->ncer; ++j) h->ja[cont + j] = b->cer[j];
    cont += b->ncer;
  }
  h->ia[h->ne] = (int)(cont + 1);
#pragma omp parallel for schedule(dynamic, 64)
  for (int m = 1; m <= h->ne; ++m) {
    for (long long k = h->ia[m - 1] - 1; k < h->ia[m] - 1; ++k) {
      cplx v = zmn(h, m, h->ja[k]);
      if (h->fp64_near)
        h->z8[k] = v;
      else
        h->z4[k] = (float)creal(v) + I * (float)cimag(v);
    }
  }
}

or_mlfma *or_mlfma_init(double lambda, double ladomax, int fp64_near, int ne,
                        const double *e_centro, const int *e_t, const int *e_po,
                        const double *e_long, int nt, const double *t_area,
                        const double *t_baric, const double *t_baric_sub, int np,
                        const double *p_coord, const int *t_p, const int *e_p,
                        const double *t_normal, int *status) {
  or_set_lambda(lambda);
  *status = 0;
  or_mlfma *h = (or_mlfma *)calloc(1, sizeof(or_mlfma));
  h->ne = ne;
  h->nt = nt;
  h->np = np;
  h->fp64_near = fp64_near;
  h->e_centro = e_centro;
  h->e_t = e_t;
  h->e_po = e_po;
  h->e_long = e_long;
  h->t_area = t_area;
  h->t_baric = t_baric;
  h->t_baric_sub = t_baric_sub;
  h->p_coord = p_coord;
  h->t_p = t_p;
  h->e_p = e_p;
  h->t_normal = t_normal;
  /* limites_dispersor (modGlobalMetod.f90:386-434) */
  int lim[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 1; i < ne; ++i) {
    const double *c = e_centro + 3 * i;
    if (c[0] > e_centro[3 * lim[0] + 0]) lim[0] = i;
    if (c[0] < e_centro[3 * lim[1] + 0]) lim[1] = i;
    if (c[1] > e_centro[3 * lim[2] + 1]) lim[2] = i;
    if (c[1] < e_centro[3 * lim[3] + 1]) lim[3] = i;
    if (c[2] > e_centro[3 * lim[4] + 2]) lim[4] = i;
    if (c[2] < e_centro[3 * lim[5] + 2]) lim[5] = i;
  }
  double mx = e_centro[3 * lim[0]], mnx = e_centro[3 * lim[1]];
  double my = e_centro[3 * lim[2] + 1], mny = e_centro[3 * lim[3] + 1];
  double mz = e_centro[3 * lim[4] + 2], mnz = e_centro[3 * lim[5] + 2];
  double dx = mx - mnx, dy = my - mny, dz = mz - mnz;
  double roo[3] = {0.5 * (mx + mnx), 0.5 * (my + mny), 0.5 * (mz + mnz)};
  double lm;
  if (dx >= dy && dx >= dz) lm = dx;
  else if (dy >= dx && dy >= dz) lm = dy;
  else lm = dz;
  double tol = g_lambda / 10000.0;
  lm = lm + tol;
  double l2f = (double)log10f(2.0f);
  int L = (int)ceil(1.0 + (log10(lm / (ladomax * g_lambda)) / l2f));
  if (L <= 3) {
    *status = 1;
    free(h);
    return NULL;
  }
  h->L = L;
  h->lv = (Level *)calloc((size_t)L + 1, sizeof(Level));
  const double sq3 = (double)sqrtf(3.0f), third = (double)(1.0f / 3.0f);
  for (int i = 1; i <= L; ++i) {
    Level *lv = &h->lv[i];
    lv->lado = lm / ldexp(1.0, i - 1);
    double d = sq3 * lv->lado;
    double kd = (2.0 * PI / g_lambda) * d;
    lv->nmult = (int)floor(kd + 6 * pow(kd, third));
    lv->nth = lv->nmult;
    lv->nph = (int)(2.0 * lv->nmult);
    lv->nk = lv->nth * lv->nph;
    if (i >= 3) fill_kappas(lv);
  }
  Box *root = add_box(&h->lv[1]);
  memcpy(root->c, roo, sizeof(roo));
  root->np = ne;
  int *all = (int *)malloc(sizeof(int) * (size_t)ne);
  for (int i = 0; i < ne; ++i) all[i] = i + 1;
  root->prop = all;
  h->rad = (cplx *)malloc(sizeof(cplx) * (size_t)h->lv[L].nk * 2 * (size_t)ne);
  geom(h, roo, lm, all, ne, 1);
  for (int i = 3; i <= L; ++i) {
    Level *lv = &h->lv[i];
    size_t sz = (size_t)lv->nk * 2 * (size_t)lv->nb;
    lv->C = (cplx *)calloc(sz, sizeof(cplx));
    lv->B = (cplx *)calloc(sz, sizeof(cplx));
    lv->A = (cplx *)calloc(sz, sizeof(cplx));
  }
  fill_transfers(h);
  fill_near(h);
  fill_interp(h);
  return h;
}

void or_mlfma_free(or_mlfma *h) {
  if (!h) return;
  for (int i = 1; i <= h->L; ++i) {
    Level *lv = &h->lv[i];
    for (int b = 0; b < lv->nb; ++b) {
      free(lv->b[b].prop);
      free(lv->b[b].cer);
      free(lv->b[b].far_box);
      free(lv->b[b].far_tr);
    }
    free(lv->b);
    free(lv->kap);
    free(lv->ath);
    free(lv->aph);
    free(lv->ang);
    free(lv->pes);
    free(lv->T);
    free(lv->rab);
    free(lv->wint);
    free(lv->iint);
    free(lv->C);
    free(lv->B);
    free(lv->A);
  }
  free(h->lv);
  free(h->rad);
  free(h->ia);
  free(h->ja);
  free(h->z4);
  free(h->z8);
  free(h->leaf);
  free(h);
}

/* ---------------------------------------------------------------- matvec */

#define FLD(lv, arr, k, p, b) (lv)->arr[((size_t)(b) * 2 + (p)) * (lv)->nk + (k)]

/* Interpolar (modMLFMA.f90:1343-1365): out(K_n,2) from in(K_{n+1},2). */
static void interpolar(const Level *lc, const cplx *in, int nkf, cplx *out) {
  for (int i = 0; i < lc->nk; ++i) {
    cplx a0 = 0, a1 = 0;
    for (int j = 0; j < 9; ++j) {
      double w = lc->wint[9 * i + j];
      int id = lc->iint[9 * i + j] - 1;
      a0 = a0 + w * in[id];
      a1 = a1 + w * in[nkf + id];
    }
    out[i] = a0;
    out[lc->nk + i] = a1;
  }
}
/* Anterpolar (modMLFMA.f90:1370-1390) */
static void anterpolar(const Level *lc, const cplx *in, int nkf, cplx *out) {
  for (int i = 0; i < 2 * nkf; ++i) out[i] = 0;
  for (int f = 0; f < lc->nk; ++f)
    for (int j = 0; j < 9; ++j) {
      double w = lc->wint[9 * f + j];
      int id = lc->iint[9 * f + j] - 1;
      out[id] = out[id] + w * in[f];
      out[nkf + id] = out[nkf + id] + w * in[lc->nk + f];
    }
}

static void far_part(or_mlfma *h, const cplx *x, cplx *res) {
  int L = h->L;
  double k = (2.0 * PI) / g_lambda;
  /* llenar_campoCyB (modMLFMA.f90:1783-1855) */
  for (int n = 3; n <= L; ++n) {
    Level *lv = &h->lv[n];
    size_t sz = (size_t)lv->nk * 2 * (size_t)lv->nb;
    memset(lv->C, 0, sz * sizeof(cplx));
    memset(lv->B, 0, sz * sizeof(cplx));
  }
  Level *ll = &h->lv[L];
  for (int c = 0; c < ll->nb; ++c)
    for (int f = 0; f < ll->b[c].np; ++f) {
      int id = ll->b[c].prop[f] - 1;
      const cplx *rf = h->rad + (size_t)id * 2 * ll->nk;
      for (int q = 0; q < 2 * ll->nk; ++q)
        ll->C[(size_t)c * 2 * ll->nk + q] = ll->C[(size_t)c * 2 * ll->nk + q] + rf[q] * x[id];
    }
  for (int n = L; n >= 4; --n) {
    Level *lv = &h->lv[n], *pl = &h->lv[n - 1];
    cplx *ag = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)pl->nk);
    for (int c = 0; c < lv->nb; ++c) {
      int par = lv->b[c].padre - 1;
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = lv->b[c].c[d] - pl->b[par].c[d];
      interpolar(pl, lv->C + (size_t)c * 2 * lv->nk, lv->nk, ag);
      for (int m = 0; m < pl->nk; ++m) {
        double kr = dot3(pl->kap + 3 * m, r);
        cplx e = cexp(I * (k * kr));
        FLD(pl, C, m, 0, par) = FLD(pl, C, m, 0, par) + ag[m] * e;
        FLD(pl, C, m, 1, par) = FLD(pl, C, m, 1, par) + ag[pl->nk + m] * e;
      }
    }
    free(ag);
  }
  for (int n = 3; n <= L; ++n) {
    Level *lv = &h->lv[n];
    for (int c = 0; c < lv->nb; ++c) {
      Box *b = &lv->b[c];
      for (int m = 0; m < b->nfar; ++m) {
        int fb = b->far_box[m] - 1, tr = b->far_tr[m] - 1;
        const cplx *T = lv->T + (size_t)tr * lv->nk;
        for (int p = 0; p < 2; ++p)
          for (int q = 0; q < lv->nk; ++q)
            FLD(lv, B, q, p, c) = FLD(lv, B, q, p, c) + FLD(lv, C, q, p, fb) * T[q];
      }
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < lv->nk; ++q) FLD(lv, B, q, p, c) = FLD(lv, B, q, p, c) * lv->pes[q];
    }
  }
  /* llenar_campoA (modMLFMA.f90:1857-1897) */
  for (int n = 4; n <= L; ++n) {
    Level *lv = &h->lv[n], *pl = &h->lv[n - 1];
    cplx *datos = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)pl->nk);
    cplx *an = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)lv->nk);
    for (int c = 0; c < lv->nb; ++c) {
      int par = lv->b[c].padre - 1;
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = lv->b[c].c[d] - pl->b[par].c[d];
      const cplx *src = (n == 4) ? pl->B : pl->A;
      for (int m = 0; m < pl->nk; ++m) {
        double kr = dot3(pl->kap + 3 * m, r);
        cplx e = cexp(I * (-(k * kr)));
        datos[m] = src[((size_t)par * 2 + 0) * pl->nk + m] * e;
        datos[pl->nk + m] = src[((size_t)par * 2 + 1) * pl->nk + m] * e;
      }
      anterpolar(pl, datos, lv->nk, an);
      for (int q = 0; q < 2 * lv->nk; ++q)
        lv->A[(size_t)c * 2 * lv->nk + q] = lv->B[(size_t)c * 2 * lv->nk + q] + an[q];
    }
    free(datos);
    free(an);
  }
  /* receive (modMLFMA.f90:1915-1933), funcREC = conjg(funcRAD) under EFIE */
  cplx cc = I * g_w * MU / (4.0 * PI);
  for (int c = 0; c < ll->nb; ++c)
    for (int f = 0; f < ll->b[c].np; ++f) {
      int id = ll->b[c].prop[f] - 1;
      const cplx *rf = h->rad + (size_t)id * 2 * ll->nk;
      const cplx *a = ll->A + (size_t)c * 2 * ll->nk;
      cplx s = 0;
      for (int q = 0; q < ll->nk; ++q) s = s + (conj(rf[q]) * a[q] + conj(rf[ll->nk + q]) * a[ll->nk + q]);
      res[id] = s * cc;
    }
}

static void near_part(or_mlfma *h, const cplx *x, cplx *res) {
  for (int m = 0; m < h->ne; ++m)
    for (long long k = h->ia[m] - 1; k < h->ia[m + 1] - 1; ++k) {
      cplx z = h->fp64_near ? h->z8[k] : (cplx)h->z4[k];
      res[m] = res[m] + x[h->ja[k] - 1] * z;
    }
}

void or_matvec(or_mlfma *h, const double *x, double *y) {
  far_part(h, (const cplx *)x, (cplx *)y);
  near_part(h, (const cplx *)x, (cplx *)y);
}
void or_matvec_far(or_mlfma *h, const double *x, double *y) {
  far_part(h, (const cplx *)x, (cplx *)y);
}
void or_matvec_near(or_mlfma *h, const double *x, double *y) {
  memset(y, 0, sizeof(cplx) * (size_t)h->ne);
  near_part(h, (const cplx *)x, (cplx *)y);
}

/* ---------------------------------------------------------------- queries */

int or_nlevels(const or_mlfma *h) { return h->L; }
int or_level_info(const or_mlfma *h, int level, int *nboxes, int *nmult,
                  int *nth, int *nph, int *ntransf, long long *npairs) {
  if (level < 1 || level > h->L) return -1;
  const Level *lv = &h->lv[level];
  *nboxes = lv->nb;
  *nmult = lv->nmult;
  *nth = lv->nth;
  *nph = lv->nph;
  *ntransf = lv->nT;
  long long p = 0;
  for (int b = 0; b < lv->nb; ++b) p += lv->b[b].nfar;
  *npairs = p;
  return 0;
}
long long or_near_nnz(const or_mlfma *h) { return h->nnz; }
void or_near_csr(const or_mlfma *h, int *ia, int *ja, double *z) {
  memcpy(ia, h->ia, sizeof(int) * (size_t)(h->ne + 1));
  memcpy(ja, h->ja, sizeof(int) * (size_t)h->nnz);
  for (long long k = 0; k < h->nnz; ++k) {
    cplx v = h->fp64_near ? h->z8[k] : (cplx)h->z4[k];
    z[2 * k] = creal(v);
    z[2 * k + 1] = cimag(v);
  }
}
long long or_level_table(const or_mlfma *h, int level, int which, double *out) {
  if (level < 3 || level > h->L) return -1;
  const Level *lv = &h->lv[level];
  long long n = 0;
  switch (which) {
    case 0: n = 3LL * lv->nk; if (out) memcpy(out, lv->kap, sizeof(double) * n); break;
    case 1: n = 2LL * lv->nk; if (out) memcpy(out, lv->ang, sizeof(double) * n); break;
    case 2: n = lv->nk; if (out) memcpy(out, lv->pes, sizeof(double) * n); break;
    case 3: if (!lv->wint) return 0; n = 9LL * lv->nk; if (out) memcpy(out, lv->wint, sizeof(double) * n); break;
    case 4: if (!lv->iint) return 0; n = 9LL * lv->nk; if (out) for (long long i = 0; i < n; ++i) out[i] = lv->iint[i]; break;
    case 5: n = 2LL * lv->nk * lv->nT; if (out) memcpy(out, lv->T, sizeof(double) * n); break;
    case 6: if (level != h->L) return 0; n = 4LL * lv->nk * h->ne; if (out) memcpy(out, h->rad, sizeof(double) * n); break;
    case 7: n = 3LL * lv->nT; if (out) memcpy(out, lv->rab, sizeof(double) * n); break;
    default: return -1;
  }
  return n;
}
void or_leaf_of_basis(const or_mlfma *h, int *leaf) {
  memcpy(leaf, h->leaf, sizeof(int) * (size_t)h->ne);
}
void or_box_centres(const or_mlfma *h, int level, double *c, int *parent) {
  const Level *lv = &h->lv[level];
  for (int b = 0; b < lv->nb; ++b) {
    memcpy(c + 3 * b, lv->b[b].c, sizeof(double) * 3);
    parent[b] = lv->b[b].padre;
  }
}

long long or_far_lists(const or_mlfma *h, int level, int *ptr, int *boxes) {
  const Level *lv = &h->lv[level];
  long long n = 0;
  if (ptr) ptr[0] = 0;
  for (int b = 0; b < lv->nb; ++b) {
    for (int m = 0; m < lv->b[b].nfar; ++m) {
      if (boxes) boxes[n] = lv->b[b].far_box[m] - 1;
      ++n;
    }
    if (ptr) ptr[b + 1] = (int)n;
  }
  return n;
}

/* ---------------------------------------------------------------- ILU0 */

int or_ilu0(int n, const int *ia, const int *ja, const double *a, int fp64,
            double *alu_out, int *jlu, int *ju) {
  const cplx *ac = (const cplx *)a;
  long long nnz = ia[n] - 1;
  cplxf *a4 = NULL;
  cplx *a8 = NULL;
  if (fp64)
    a8 = (cplx *)calloc((size_t)nnz + 2, sizeof(cplx));
  else
    a4 = (cplxf *)calloc((size_t)nnz + 2, sizeof(cplxf));
  /* 1-based views */
#define ALU_GET(k) (fp64 ? a8[(k)] : (cplx)a4[(k)])
#define ALU_SET(k, v)                                                   \
  do {                                                                  \
    cplx _v = (v);                                                      \
    if (fp64) a8[(k)] = _v;                                             \
    else a4[(k)] = (float)creal(_v) + I * (float)cimag(_v);             \
  } while (0)
  int *iw = (int *)calloc((size_t)n + 1, sizeof(int));
  int ju0 = n + 2;
  jlu[0] = ju0; /* jlu(1) */
  int ierr = 0;
  for (int ii = 1; ii <= n; ++ii) {
    int js = ju0;
    for (int j = ia[ii - 1]; j <= ia[ii] - 1; ++j) {
      int jcol = ja[j - 1];
      cplx aj = ac[j - 1];
      if (!fp64) aj = (float)creal(aj) + I * (float)cimag(aj);
      if (jcol == ii) {
        ALU_SET(ii, aj);
        iw[jcol] = ii;
        ju[ii - 1] = ju0;
      } else {
        ALU_SET(ju0, aj);
        jlu[ju0 - 1] = jcol;
        iw[jcol] = ju0;
        ju0 = ju0 + 1;
      }
    }
    jlu[ii] = ju0; /* jlu(ii+1) */
    int jf = ju0 - 1;
    int jm = ju[ii - 1] - 1;
    for (int j = js; j <= jm; ++j) {
      int jrow = jlu[j - 1];
      cplx tl;
      if (fp64) {
        tl = a8[j] * a8[jrow];
      } else {
        /* complex(4) * complex(4) evaluated in single precision */
        float ar = crealf(a4[j]), ai = cimagf(a4[j]);
        float br = crealf(a4[jrow]), bi = cimagf(a4[jrow]);
        float tr = ar * br - ai * bi, ti = ar * bi + ai * br;
        tl = (double)tr + I * (double)ti;
      }
      ALU_SET(j, tl);
      for (int jj = ju[jrow - 1]; jj <= jlu[jrow] - 1; ++jj) {
        int jw = iw[jlu[jj - 1]];
        if (jw != 0) ALU_SET(jw, ALU_GET(jw) - tl * ALU_GET(jj));
      }
    }
    cplx d = ALU_GET(ii);
    if (creal(d) == 0.0 && cimag(d) == 0.0) {
      ierr = ii;
      break;
    }
    ALU_SET(ii, 1.0 / d);
    iw[ii] = 0;
    for (int i = js; i <= jf; ++i) iw[jlu[i - 1]] = 0;
  }
  for (long long k = 1; k <= nnz + 1; ++k) {
    cplx v = ALU_GET(k);
    alu_out[2 * (k - 1)] = creal(v);
    alu_out[2 * (k - 1) + 1] = cimag(v);
  }
#undef ALU_GET
#undef ALU_SET
  free(iw);
  free(a4);
  free(a8);
  return ierr;
}

void or_lusol0(int n, const double *yv, double *xv, const double *alu_d,
               const int *jlu, const int *ju) {
  const cplx *y = (const cplx *)yv;
  const cplx *alu = (const cplx *)alu_d; /* alu(k) at index k-1 */
  cplx *x = (cplx *)xv;
  for (int i = 1; i <= n; ++i) {
    cplx s = y[i - 1];
    for (int k = jlu[i - 1]; k <= ju[i - 1] - 1; ++k) s = s - alu[k - 1] * x[jlu[k - 1] - 1];
    x[i - 1] = s;
  }
  for (int i = n; i >= 1; --i) {
    cplx s = x[i - 1];
    for (int k = ju[i - 1]; k <= jlu[i] - 1; ++k) s = s - alu[k - 1] * x[jlu[k - 1] - 1];
    x[i - 1] = alu[i - 1] * s;
  }
}

/* ---------------------------------------------------------------- CGS */

static double cnorm(const cplx *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += creal(conj(v[i]) * v[i]);
  return sqrt(s);
}

int or_cgs(or_mlfma *h, const double *bv, const double *gv, double errmax,
           const double *alu, const int *jlu, const int *ju, int maxit,
           double *xv, int *converged, double *relres) {
  int n = h->ne;
  const cplx *b = (const cplx *)bv, *guess = (const cplx *)gv;
  cplx *x = (cplx *)xv;
  size_t sz = sizeof(cplx) * (size_t)n;
  cplx *r = malloc(sz), *R_ = malloc(sz), *u = malloc(sz), *p = malloc(sz), *P_ = malloc(sz),
       *V_ = malloc(sz), *q = malloc(sz), *U_ = malloc(sz), *Q_ = malloc(sz), *t = malloc(sz);
  cplx rho1 = 0, rho2 = 0;
  int prec = alu && jlu && ju;
  or_matvec(h, (const double *)guess, (double *)t);
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - t[i];
    R_[i] = r[i];
    x[i] = guess[i];
  }
  double nb = cnorm(b, n);
  int it = 0;
  *converged = 0;
  *relres = cnorm(r, n) / nb;
  while (it < maxit) {
    ++it;
    rho1 = 0;
    for (int i = 0; i < n; ++i) rho1 = rho1 + r[i] * R_[i];
    if (rho1 == 0) break;
    if (it == 1) {
      for (int i = 0; i < n; ++i) { u[i] = r[i]; p[i] = u[i]; }
    } else {
      cplx beta = rho1 / rho2;
      for (int i = 0; i < n; ++i) {
        u[i] = r[i] + beta * q[i];
        p[i] = u[i] + beta * (q[i] + beta * p[i]);
      }
    }
    if (prec) or_lusol0(n, (double *)p, (double *)P_, alu, jlu, ju);
    else memcpy(P_, p, sz);
    or_matvec(h, (const double *)P_, (double *)V_);
    cplx den = 0;
    for (int i = 0; i < n; ++i) den = den + R_[i] * V_[i];
    cplx alfa = rho1 / den;
    for (int i = 0; i < n; ++i) {
      q[i] = u[i] - alfa * V_[i];
      t[i] = u[i] + q[i];
    }
    if (prec) or_lusol0(n, (double *)t, (double *)U_, alu, jlu, ju);
    else memcpy(U_, t, sz);
    for (int i = 0; i < n; ++i) x[i] = x[i] + alfa * U_[i];
    or_matvec(h, (const double *)U_, (double *)Q_);
    for (int i = 0; i < n; ++i) r[i] = r[i] - alfa * Q_[i];
    double nr = cnorm(r, n);
    *relres = nr / nb;
    if (nr < errmax * nb) {
      *converged = 1;
      break;
    }
    rho2 = rho1;
  }
  free(r); free(R_); free(u); free(p); free(P_); free(V_); free(q); free(U_); free(Q_); free(t);
  return it;
}

/* ---------------------------------------------------------------- excitation, RCS */

void or_plane_wave(or_mlfma *h, const double *d, const double *ctte_,
                   const double *pol_, double *bv) {
  cplx ctte = ctte_[0] + I * ctte_[1];
  cplx pol[3] = {pol_[0] + I * pol_[1], pol_[2] + I * pol_[3], pol_[4] + I * pol_[5]};
  double nd = sqrt(dot3(d, d));
  double U[3] = {d[0] / nd, d[1] / nd, d[2] / nd};
  double k = g_w * sqrt(MU * EPSI);
  cplx *b = (cplx *)bv;
  for (int i = 0; i < h->ne; ++i) {
    const double *bp = h->t_baric + 3 * (h->e_t[2 * i] - 1);
    const double *bm = h->t_baric + 3 * (h->e_t[2 * i + 1] - 1);
    const double *vp = h->p_coord + 3 * (h->e_po[2 * i] - 1);
    const double *vm = h->p_coord + 3 * (h->e_po[2 * i + 1] - 1);
    cplx ep = cexp(I * (-(k * dot3(U, bp))));
    cplx em = cexp(I * (-(k * dot3(U, bm))));
    cplx s1 = 0, s2 = 0;
    for (int c = 0; c < 3; ++c) {
      s1 = s1 + (pol[c] * ctte * ep) * (bp[c] - vp[c]);
      s2 = s2 + (pol[c] * ctte * em) * (vm[c] - bm[c]);
    }
    b[i] = (s1 + s2) * 0.5 * h->e_long[i];
  }
}

double or_sigma_rcs(or_mlfma *h, const double *Iv, const double *pol_,
                    const double *ctte_, double theta, double phi) {
  const cplx *Ic = (const cplx *)Iv;
  cplx ctte = ctte_[0] + I * ctte_[1];
  cplx pol[3] = {pol_[0] + I * pol_[1], pol_[2] + I * pol_[3], pol_[4] + I * pol_[5]};
  /* calc_RCS (modRCS.f90:109-153) */
  const double *ec = h->e_centro;
  int lim[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 1; i < h->ne; ++i) {
    const double *c = ec + 3 * i;
    if (c[0] > ec[3 * lim[0] + 0]) lim[0] = i;
    if (c[0] < ec[3 * lim[1] + 0]) lim[1] = i;
    if (c[1] > ec[3 * lim[2] + 1]) lim[2] = i;
    if (c[1] < ec[3 * lim[3] + 1]) lim[3] = i;
    if (c[2] > ec[3 * lim[4] + 2]) lim[4] = i;
    if (c[2] < ec[3 * lim[5] + 2]) lim[5] = i;
  }
  double mx = ec[3 * lim[0]], mnx = ec[3 * lim[1]], my = ec[3 * lim[2] + 1],
         mny = ec[3 * lim[3] + 1], mz = ec[3 * lim[4] + 2], mnz = ec[3 * lim[5] + 2];
  double lam = 2 * PI / (g_w * sqrt(MU * EPSI));
  double D = sqrt(((mx - mnx) * (mx - mnx) + (my - mny) * (my - mny)) + (mz - mnz) * (mz - mnz));
  double rzl = 2 * (D * D) / lam;
  if (50 * D > rzl) rzl = 50 * D;
  if (20 * lam > rzl) rzl = 20 * lam;
  rzl = rzl * 10;
  double roo[3] = {0.5 * (mx + mnx), 0.5 * (my + mny), 0.5 * (mz + mnz)};
  double ro[3] = {roo[0] + (rzl * sin(theta) * cos(phi)), roo[1] + (rzl * sin(theta) * sin(phi)),
                  roo[2] + (rzl * cos(theta))};
  /* calc_E (modRCS.f90:60-89) */
  cplx E[3] = {0, 0, 0};
  for (int i = 0; i < h->ne; ++i) {
    const double *t1 = h->t_baric + 3 * (h->e_t[2 * i] - 1);
    const double *t2 = h->t_baric + 3 * (h->e_t[2 * i + 1] - 1);
    double rm[3], rrm[3];
    cplx m[3];
    for (int c = 0; c < 3; ++c) {
      rm[c] = 0.5 * (t1[c] + t2[c]);
      m[c] = (t2[c] - t1[c]) * Ic[i] * h->e_long[i];
      rrm[c] = ro[c] - rm[c];
    }
    cplx sm = (m[0] * rrm[0] + m[1] * rrm[1]) + m[2] * rrm[2];
    double rr = dot3(rrm, rrm);
    cplx g = green(ro, rm);
    for (int c = 0; c < 3; ++c) E[c] = E[c] + g * (sm * rrm[c] / rr - m[c]);
  }
  cplx cc = I * g_w * MU / (4 * PI);
  for (int c = 0; c < 3; ++c) E[c] = E[c] * cc;
  double s = creal(conj(E[0]) * E[0] + conj(E[1]) * E[1] + conj(E[2]) * E[2]);
  cplx pp[3] = {pol[0] * ctte, pol[1] * ctte, pol[2] * ctte};
  double d = creal(conj(pp[0]) * pp[0] + conj(pp[1]) * pp[1] + conj(pp[2]) * pp[2]);
  return s * 4 * PI * (rzl * rzl) / d;
}
