/*
 * mlegs_oracle.c -- CPU FP64 restatement of the MLegS pseudo-spectral
 * Boussinesq time step (reference: Jinge-Wang/MLegS, Fortran 90).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for libmlegs_hip.so and the
 * CPU baseline timed beside it.  The product path never links or calls it.
 *
 * Each function names the reference routine (file:line) whose semantics it
 * restates.  Loops are written from the mathematics and the reference's
 * observable behaviour (chop state, log-term bookkeeping, summation
 * structure); floating-point evaluation order follows the reference
 * expressions where that is visible in the source.
 *
 * Parity unpinned: SPESLOP's overall scale.  Its structure is pinned (odd
 * and linear in log E, clamps, floors), but the factor a pure power law's
 * slope comes out with (about 4.55 at 120 modes) is the reference's own
 * discretisation and has no check independent of the formula.  DIAGNOST's
 * VALANG (INTEG_P's storage-order read of FFF coefficients) and the
 * CALCAT1 spreads (the pinned AT1 table summed against the m = 0 column)
 * have no further check either.
 */
#include "mlegs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

enum {
  E_OK = 0, E_SCALAR3 = 1, E_NAN = 10, E_DEALIAS = 11, E_DIVERGE = 12, E_LOGTERM = 13,
  E_TRANSFORM = 2, E_LEGOP = 3, E_CALC = 31, E_DIM = 32, E_SOLVER = 5
};

struct og_grid {
  og_params p;
  int nr, nrh, nth, nx, ntchop, nxchop, minc, mklink;
  double ell, ell2, zlen, zlen0;
  int ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim, nxchoph;
  int nrchop;
  int *nrchops;          /* [ntchopdim] */
  int *m;                /* [ntchopdim] */
  double *ak;            /* [ntchopdim * nxchopdim], (mm,kk) col-major */
  double *r, *x, *w, *lnx;
  int nnorm;             /* nrchopdim + 14 */
  double *norm, *lognorm;/* [nnorm * ntchopdim] */
  int npf;               /* nrchopdim + 1 */
  double *pf;            /* [nrh * npf * ntchopdim] */
  double *at0, *at1;     /* [nrchopdim] */
  double *thr, *thi, *z;
  /* exponential hyperviscosity filters (HYP3SET) */
  double *hv_r, *hv_t, *hv_x, *hd_r, *hd_t, *hd_x;
  /* HYP3ADJ's directional coefficients and controller history (VISC%,
   * mod_march.f90:75-76) */
  double nup_r, nup_t, nup_x, kap_r, kap_t, kap_x;
  double nupb, curvmo, curvko, slopmo, slopko;
  double kappapb, pcurvmo, pcurvko, pslopmo, pslopko;
  /* axial sponge (RMV%SW = 2): TOPHAT_X over the NX z points, tau */
  int rmv_sw;
  double *tophat_x, sp_tau;
};

struct og_solver {
  og_grid *g;
  double t;
  int n;
  og_field *n_psi, *n_chi, *n_b;                       /* AB2 history */
  og_field *r_n_r, *r_n_p, *n_z, *r_pi_r, *r_pi_p, *pi_z; /* ETD history */
  double etd_e[16], etd_nl[16];
  /* TIME_STEPPING's REMOVE and HYP3ADJ cadence (RMV%SW/INT, VISC%ADJSW/ADJINT) */
  int rmv_sw, rmv_int, adj_sw, adj_int;
};

#define FIDX(g, n, mm, kk) ((size_t)(n) + (size_t)(g)->nrchopdim * ((size_t)(mm) + (size_t)(g)->ntchopdim * (size_t)(kk)))
#define QIDX(g, r, mm, kk) ((size_t)(r) + (size_t)(g)->ndimr * ((size_t)(mm) + (size_t)(g)->ntchopdim * (size_t)(kk)))
#define PIDX(g, r, j, z) ((size_t)(r) + (size_t)(g)->ndimr * ((size_t)(j) + (size_t)(g)->ndimth * (size_t)(z)))
#define PF(g, i, n, mm) ((g)->pf[(size_t)(i) + (size_t)(g)->nrh * ((size_t)(n) + (size_t)(g)->npf * (size_t)(mm))])
#define LOGNORM(g, n, mm) ((g)->lognorm[(size_t)(n) + (size_t)(g)->nnorm * (size_t)(mm)])
#define NORM(g, n, mm) ((g)->norm[(size_t)(n) + (size_t)(g)->nnorm * (size_t)(mm)])
#define AK(g, mm, kk) ((g)->ak[(size_t)(mm) + (size_t)(g)->ntchopdim * (size_t)(kk)])

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* Summation order of the Legendre sums (RTRAN both ways, PROJECT):
 * 0 ascending as in the reference, 1 descending.  Test use only: measures
 * how much a result depends on the order of its floating-point sums. */
static int sum_rev = 0;
void og_set_sum_order(int reverse) { sum_rev = reverse ? 1 : 0; }

/* Test-only planted mistakes (0: none).  Each value breaks one sign or index
 * of the restatement, so a test can show that it detects that mistake:
 *  1 buoyancy force +b z (FORCE)      2 Coriolis theta component +2 Omega u_r
 *  3 N^2 u_z term sign (FORCE)        4 sign of KE's chi term
 *  5 sign of DKE                      6 sign of PC2VOR's i m del^2 chi term
 *  7 sign of GRAD's i m b             8 ETD k = 0 row replacement left out
 *  9 CHOPDO keeps one row too many    10 sign of GRAD's i k b */
static int mutation = 0;
void og_set_mutation(int k) { mutation = k; }

void og_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ===================================================================== */
/* Gauss-Legendre nodes and weights: mod_lin_legendre.f90:79-131         */
/* Newton iteration from the Numerical-Recipes start, stopped when two    */
/* successive iterates differ by <= 1e-15; the weight uses the last       */
/* derivative evaluation (at the previous iterate), as the reference does.*/
/* ===================================================================== */
static void gauss_legendre(int n, double *x, double *w) {
  const double pi = acos(-1.0);
  int mh = (int)ceil((n + 1) / 2.0);
  for (int i = 1; i <= mh; i++) {
    double z = cos(pi * (i - 0.25) / (n + 0.5));
    double z1 = 0.0, pp = 0.0;
    do {
      double p1 = 1.0, p2 = 0.0, p3;
      for (int j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > 1e-15);
    x[i - 1] = -z;
    x[n - i] = z;
    w[i - 1] = 2.0 / ((1 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* LEG_LOG_NORMT: mod_lin_legendre.f90:306-377.  0.5*log((2n+1)/2 (n-m)!/(n+m)!) */
static void leg_log_norm(int ndim, int m, double *out) {
  double wk = log(0.5);
  for (int mm = 1; mm <= m; mm++)
    wk = wk + log(2 * mm + 1.0) - log((double)(2 * mm) * ((2 * mm - 1.0) * (2 * mm - 1.0)));
  out[0] = wk;
  for (int nn = 2; nn <= ndim; nn++) {
    int n = m + nn - 1;
    out[nn - 1] = out[nn - 2] + log(2 * n + 1.0) - log(2 * n - 1.0) + log(1.0 * (n - m)) - log(1.0 * (n + m));
  }
  for (int i = 0; i < ndim; i++) out[i] = 0.5 * out[i];
}

/* LEG_NORMT: mod_lin_legendre.f90:205-278 (plain normalisation factors). */
static void leg_norm(int ndim, int m, double *out) {
  double wk = 0.5;
  for (int mm = 1; mm <= m; mm++) wk = wk * (2 * mm + 1.0) / ((double)(2 * mm) * ((2 * mm - 1.0) * (2 * mm - 1.0)));
  out[0] = wk * pow(100.0, m);
  for (int nn = 2; nn <= ndim; nn++) {
    int n = m + nn - 1;
    out[nn - 1] = out[nn - 2] * (2 * n + 1.0) / (2 * n - 1.0) * (n - m) / (n + m);
  }
  for (int i = 0; i < ndim; i++) out[i] = sqrt(out[i]) * pow(10.0, -m);
}

/* LOGFACT: mod_lin_legendre.f90:707-725, log((2n-1)!!) in FP64. */
static double logfact(int n) { return lgamma(2 * n + 1.0) - n * log(2.0) - lgamma(n + 1.0); }

/* LEG_TBL3: mod_lin_legendre.f90:526-644.  Recurrence on P_n^m/(2n-1)!!
 * carried in extended precision (80-bit long double instead of the
 * reference's 50-digit FM), scaled by exp(lognorm + logfact) formed from the
 * FP64 logs, rounded to FP64; results below the FP64 normal range -> 0. */
static void leg_tbl(const double *x, int nx, int ne, int m, const double *lognorm, double *out /* [nx*ne] */) {
  long double *q = (long double *)malloc(sizeof(long double) * (size_t)ne);
  for (int xx = 0; xx < nx; xx++) {
    long double xl = (long double)x[xx];
    if (m == 0) {
      q[0] = 1.0L;
    } else {
      long double s = sqrtl(1.0L - xl * xl);
      long double v = (m % 2) ? -1.0L : 1.0L;
      for (int i = 0; i < m; i++) v *= s;
      q[0] = v;
    }
    if (ne > 1) q[1] = xl * q[0];
    for (int nn = 3; nn <= ne; nn++) {
      int n = m + nn - 1;
      q[nn - 1] = 1.0L / (n - m) * (q[nn - 2] * xl - q[nn - 3] * (long double)(n + m - 1) / (long double)(2 * n - 1) / (long double)(2 * n - 3));
    }
    for (int nn = 1; nn <= ne; nn++) {
      long double e = (long double)lognorm[nn - 1] + (long double)logfact(m + nn - 1);
      long double v = q[nn - 1] * expl(e);
      double d = (double)v;
      if (fabs(d) < DBL_MIN) d = 0.0;
      out[xx + (size_t)nx * (nn - 1)] = d;
    }
  }
  free(q);
}

/* ===================================================================== */
/* Band operators (normalised basis): mod_lin_legendre.f90               */
/* Band storage B(row, b) = matrix(row, row + b - center), col-major.    */
/* ===================================================================== */
/* BAND_LOGLEG_XM: mod_lin_legendre.f90:910-982, (1-x)*  [ni x 3] */
static void band_xm(int ni, int m, const double *ln, double *b) {
  memset(b, 0, sizeof(double) * 3 * (size_t)ni);
  for (int nn = 1; nn <= ni; nn++) {
    int n = m + nn - 1;
    if (nn - 1 >= 1) b[nn - 1] = -((double)(n - m)) / (2 * n - 1.0);
    b[nn - 1 + ni] = 1.0;
    if (nn + 1 <= ni) b[nn - 1 + 2 * ni] = -(n + m + 1.0) / (2 * n + 3.0);
  }
  if (ni <= 1) return;
  for (int i = 1; i < ni; i++) b[i] *= exp(ln[i - 1] - ln[i]);
  for (int i = 0; i < ni - 1; i++) b[i + 2 * ni] *= exp(ln[i + 1] - ln[i]);
}

/* BAND_LOGLEG_XXDX: mod_lin_legendre.f90:1683-1758, (1-x^2) d/dx  [ni x 3] */
static void band_xxdx(int ni, int m, const double *ln, double *b) {
  memset(b, 0, sizeof(double) * 3 * (size_t)ni);
  for (int nn = 1; nn <= ni; nn++) {
    int n = m + nn - 1;
    if (nn - 1 >= 1) b[nn - 1] = -(n - 1.0) * (n - m) / (2 * n - 1.0);
    if (nn + 1 <= ni) b[nn - 1 + 2 * ni] = (n + 2.0) * (n + m + 1.0) / (2 * n + 3.0);
  }
  if (ni <= 1) return;
  for (int i = 1; i < ni; i++) b[i] *= exp(ln[i - 1] - ln[i]);
  for (int i = 0; i < ni - 1; i++) b[i + 2 * ni] *= exp(ln[i + 1] - ln[i]);
}

/* BAND_LOGLEG_RAT_DEL2H: mod_lin_legendre.f90:1986-2092, L^2-scaled del^2_h with S=0 [ni x 5] */
static void band_del2h(int ni, int m, double ell, const double *ln, double *b) {
  const double s = 0.0;
  double ell2 = pow(ell, 2.0);
  memset(b, 0, sizeof(double) * 5 * (size_t)ni);
  for (int nn = 1; nn <= ni; nn++) {
    int n = m + nn - 1;
    if (nn - 2 >= 1) b[nn - 1] = -(n - m - 1.0) * (n - m) * (n - 2 + s) * (n - 1 + s) / (2 * n - 3.0) / (2 * n - 1.0) / ell2;
    if (nn - 1 >= 1) b[nn - 1 + ni] = 2.0 * n * (n - m) * (n - 1 + s) / (2 * n - 1.0) / ell2;
    b[nn - 1 + 2 * ni] = (-2.0 * n * (n + 1.0) * (3.0 * n * n + 3 * n - m * m - 2) + 2 * s * (s - 2) * (n * n + n + m * m - 1.0)) /
                         (2 * n - 1.0) / (2 * n + 3.0) / ell2;
    if (nn + 1 <= ni) b[nn - 1 + 3 * ni] = 2.0 * (n + 1) * (n + m + 1.0) * (n + 2 - s) / (2 * n + 3.0) / ell2;
    if (nn + 2 <= ni) b[nn - 1 + 4 * ni] = -(n + m + 1.0) * (n + m + 2.0) * (n + 3 - s) * (n + 2 - s) / (2 * n + 3.0) / (2 * n + 5.0) / ell2;
  }
  if (ni <= 1) return;
  for (int i = 1; i < ni; i++) b[i + ni] *= exp(ln[i - 1] - ln[i]);
  for (int i = 0; i < ni - 1; i++) b[i + 3 * ni] *= exp(ln[i + 1] - ln[i]);
  if (ni <= 2) return;
  for (int i = 2; i < ni; i++) b[i] *= exp(ln[i - 2] - ln[i]);
  for (int i = 0; i < ni - 2; i++) b[i + 4 * ni] *= exp(ln[i + 2] - ln[i]);
}

/* BANMULC: mod_bandmat.f90:1199-1243 (band x complex vector). */
static void banmul(const double *a, int lda, int nb, int nc, const cplx *x, int n, cplx *out) {
  for (int c = 0; c < n; c++) out[c] = a[c + (size_t)lda * nc] * x[c];
  for (int bi = 0; bi < nb; bi++) {
    if (bi == nc) continue;
    int is = bi - nc;
    int c0 = imax(0, -is), c1 = imin(n - 1, n - 1 - is);
    for (int c = c0; c <= c1; c++) out[c] += a[c + (size_t)lda * bi] * x[c + is];
  }
}

/* BANBANMUL: mod_bandmat.f90:1288-1351 (band x band, n x n truncated). */
static void banbanmul(const double *a, int aw, int ac, const double *b, int bw, int bc, int n, double *c /* [n*(aw+bw-1)] */) {
  int cw = aw + bw - 1;
  memset(c, 0, sizeof(double) * (size_t)n * cw);
  for (int i = 0; i < aw; i++) {
    int as = i - ac;
    int a0 = imax(0, -as), a1 = imin(n - 1, n - 1 - as);
    for (int j = 0; j < bw; j++) {
      int k = i + j;
      int bs = j - bc;
      int b0 = imax(0, -bs), b1 = imin(n - 1, n - 1 - bs);
      int c0 = imax(a0, b0 - as), c1 = imin(a1, b1 - as);
      for (int r = c0; r <= c1; r++) c[r + (size_t)n * k] += a[r + (size_t)n * i] * b[r + as + (size_t)n * j];
    }
  }
}

/* LUDCMPB for one system: mod_bandmat.f90:65-156.  No pivoting; stores the
 * unit-upper factor, the scaled multipliers and 1/diagonal in place. */
static void lub(double *a, int lda, int nd, int ne, int c0) {
  for (int nn = 0; nn < nd; nn++) {
    for (int b = c0 + 1; b < ne; b++)
      if (nn + (b - c0) < nd) a[nn + (size_t)lda * b] /= a[nn + (size_t)lda * c0];
    for (int b = c0 + 1; b < ne; b++)
      for (int j = 1; j <= c0; j++)
        if (nn + imax(b - c0, j) < nd)
          a[nn + j + (size_t)lda * (b - j)] -= a[nn + (size_t)lda * b] * a[nn + j + (size_t)lda * (c0 - j)];
    for (int j = 1; j <= c0; j++)
      if (nn + j < nd) a[nn + j + (size_t)lda * (c0 - j)] /= a[nn + (size_t)lda * c0];
  }
  for (int nn = 0; nn < nd; nn++) a[nn + (size_t)lda * c0] = 1.0 / a[nn + (size_t)lda * c0];
}

/* LUBKSBB for one system: mod_bandmat.f90:320-402. */
static void solveb(const double *a, int lda, int nd, int ne, int c0, cplx *x) {
  for (int nn = 1; nn < nd; nn++) {
    int nne = imin(nn, c0);
    for (int j = 1; j <= nne; j++) x[nn] -= x[nn - j] * a[nn + (size_t)lda * (c0 - j)];
  }
  for (int nn = 0; nn < nd; nn++) x[nn] *= a[nn + (size_t)lda * c0];
  for (int nn = nd - 2; nn >= 0; nn--)
    for (int i = 1; i <= ne - 1 - c0; i++)
      if (nn + i < nd) x[nn] -= x[nn + i] * a[nn + (size_t)lda * (c0 + i)];
}

/* ===================================================================== */
/* FFT (radix-2 iterative; naive DFT for other lengths)                   */
/* sign = -1: sum x_n e^{-2 pi i k n / N}; sign = +1: e^{+...}; unscaled. */
/* ===================================================================== */
static void fft_c(cplx *a, int n, int sign) {
  if (n <= 1) return;
  if ((n & (n - 1)) == 0) {
    for (int i = 1, j = 0; i < n; i++) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    const double pi = acos(-1.0);
    for (int len = 2; len <= n; len <<= 1) {
      int h = len >> 1;
      for (int j = 0; j < h; j++) {
        double ang = sign * 2.0 * pi * j / len;
        cplx wj = cos(ang) + I * sin(ang);
        for (int i = 0; i < n; i += len) {
          cplx u = a[i + j], v = a[i + j + h] * wj;
          a[i + j] = u + v;
          a[i + j + h] = u - v;
        }
      }
    }
  } else {
    cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    const double pi = acos(-1.0);
    for (int k = 0; k < n; k++) {
      cplx s = 0;
      for (int j = 0; j < n; j++) {
        long p = ((long)k * j) % n;
        double ang = sign * 2.0 * pi * (double)p / n;
        s += a[j] * (cos(ang) + I * sin(ang));
      }
      t[k] = s;
    }
    memcpy(a, t, sizeof(cplx) * (size_t)n);
    free(t);
  }
}

/* ===================================================================== */
/* Grid set-up: LEGINIT, mod_fft.f90:180-367                             */
/* ===================================================================== */
static void hyp3_init(og_grid *g);

og_grid *og_init(const og_params *p) {
  og_grid *g = (og_grid *)calloc(1, sizeof(og_grid));
  g->p = *p;
  g->nr = p->nr;
  g->nrh = p->nr / 2;
  g->nth = p->nth;
  g->nx = p->nx;
  g->ntchop = p->ntchop;
  g->nxchop = p->nxchop;
  g->minc = p->minc > 0 ? p->minc : 1;
  g->mklink = p->mklink;
  g->ell = p->ell;
  g->ell2 = pow(p->ell, 2.0);
  g->zlen = p->zlen;
  g->zlen0 = p->zlen;
  g->ndimr = g->nr + 3;
  g->ndimth = g->nth + 1;
  if (g->ndimth == 2) g->ndimth = 1;
  g->ndimx = g->nx;
  g->nrchopdim = imin(p->nrchop + 6, g->nr);
  g->ntchopdim = imin(p->ntchop, g->nth);
  g->nxchopdim = p->nxchop * 2 - 1;
  g->nxchoph = g->nx - p->nxchop + 2;
  if (p->nxchop == 1) g->nxchoph = 1;
  if ((2 * p->nxchop > g->nx) && (g->nx > 1 || p->nxchop > 1)) {
    free(g);
    return NULL;
  }
  g->nrchop = p->nrchop;
  g->m = (int *)calloc((size_t)g->ntchopdim, sizeof(int));
  for (int mm = 0; mm < g->ntchopdim; mm++) g->m[mm] = g->minc * mm;
  g->nrchops = (int *)calloc((size_t)g->ntchopdim, sizeof(int));
  og_chopset(g, 0);
  g->ak = (double *)calloc((size_t)g->ntchopdim * g->nxchopdim, sizeof(double));
  const double pi = acos(-1.0);
  for (int mm = 0; mm < g->ntchopdim; mm++)
    for (int kk = 1; kk <= g->nxchopdim; kk++) {
      int kv = kk - 1 + g->mklink * mm;
      if (kk > g->nxchop) kv = -(g->nxchopdim - kk + 1);
      AK(g, mm, kk - 1) = 2 * pi / g->zlen * kv;
    }
  int nr = g->nr;
  g->r = (double *)calloc((size_t)nr, sizeof(double));
  g->x = (double *)calloc((size_t)nr, sizeof(double));
  g->w = (double *)calloc((size_t)nr, sizeof(double));
  g->lnx = (double *)calloc((size_t)nr, sizeof(double));
  gauss_legendre(nr, g->x, g->w);
  for (int i = 0; i < nr; i++) {
    g->r[i] = g->ell * sqrt((1 + g->x[i]) / (1 - g->x[i]));
    g->lnx[i] = -log(1 - g->x[i]);
  }
  g->nnorm = g->nrchopdim + 14;
  g->norm = (double *)calloc((size_t)g->nnorm * g->ntchopdim, sizeof(double));
  g->lognorm = (double *)calloc((size_t)g->nnorm * g->ntchopdim, sizeof(double));
  for (int mm = 0; mm < g->ntchopdim; mm++) {
    leg_norm(g->nnorm, g->m[mm], &NORM(g, 0, mm));
    leg_log_norm(g->nnorm, g->m[mm], &LOGNORM(g, 0, mm));
  }
  g->npf = g->nrchopdim + 1;
  g->pf = (double *)calloc((size_t)g->nrh * g->npf * g->ntchopdim, sizeof(double));
#pragma omp parallel for schedule(dynamic)
  for (int mm = 0; mm < g->ntchopdim; mm++)
    leg_tbl(g->x, g->nrh, g->npf, g->m[mm], &LOGNORM(g, 0, mm), &PF(g, 0, 0, mm));
  g->at0 = (double *)calloc((size_t)g->nrchopdim, sizeof(double));
  g->at1 = (double *)calloc((size_t)g->nrchopdim, sizeof(double));
  double xm1 = -1.0, xp1 = 1.0;
  leg_tbl(&xm1, 1, g->nrchopdim, 0, &LOGNORM(g, 0, 0), g->at0);
  leg_tbl(&xp1, 1, g->nrchopdim, 0, &LOGNORM(g, 0, 0), g->at1);
  g->z = (double *)calloc((size_t)g->nx + 1, sizeof(double));
  for (int kk = 0; kk <= g->nx; kk++) g->z[kk] = g->zlen * kk / g->nx;
  int nt = g->minc * g->nth + 1;
  g->thr = (double *)calloc((size_t)nt, sizeof(double));
  g->thi = (double *)calloc((size_t)nt, sizeof(double));
  for (int i = 0; i < nt; i++) {
    g->thr[i] = 2 * pi / g->minc / g->nth * i;
    g->thi[i] = 2 * pi / g->minc / g->nth * (i + 0.5);
  }
  g->nup_r = g->nup_t = g->nup_x = p->nup;
  g->kap_r = g->kap_t = g->kap_x = p->kappap;
  g->nupb = -1.0;
  g->kappapb = -1.0;
  if (p->visc_sw == 3 || p->hyper_cn) hyp3_init(g);
  return g;
}

void og_free(og_grid *g) {
  if (!g) return;
  free(g->nrchops); free(g->m); free(g->ak); free(g->r); free(g->x); free(g->w); free(g->lnx);
  free(g->norm); free(g->lognorm); free(g->pf); free(g->at0); free(g->at1);
  free(g->z); free(g->thr); free(g->thi);
  free(g->hv_r); free(g->hv_t); free(g->hv_x); free(g->hd_r); free(g->hd_t); free(g->hd_x);
  free(g->tophat_x);
  free(g);
}

void og_dims(const og_grid *g, int *o) {
  o[0] = g->ndimr; o[1] = g->ndimth; o[2] = g->ndimx;
  o[3] = g->nrchopdim; o[4] = g->ntchopdim; o[5] = g->nxchopdim;
  o[6] = g->nrh; o[7] = g->nxchoph; o[8] = g->nrchop;
  o[9] = g->npf; o[10] = g->nnorm; o[11] = g->nr;
}

const double *og_table(const og_grid *g, const char *name, int *n) {
  if (!strcmp(name, "x")) { *n = g->nr; return g->x; }
  if (!strcmp(name, "w")) { *n = g->nr; return g->w; }
  if (!strcmp(name, "r")) { *n = g->nr; return g->r; }
  if (!strcmp(name, "lnx")) { *n = g->nr; return g->lnx; }
  if (!strcmp(name, "pf")) { *n = g->nrh * g->npf * g->ntchopdim; return g->pf; }
  if (!strcmp(name, "lognorm")) { *n = g->nnorm * g->ntchopdim; return g->lognorm; }
  if (!strcmp(name, "norm")) { *n = g->nnorm * g->ntchopdim; return g->norm; }
  if (!strcmp(name, "at0")) { *n = g->nrchopdim; return g->at0; }
  if (!strcmp(name, "at1")) { *n = g->nrchopdim; return g->at1; }
  if (!strcmp(name, "ak")) { *n = g->ntchopdim * g->nxchopdim; return g->ak; }
  if (!strcmp(name, "z")) { *n = g->nx + 1; return g->z; }
  if (!strcmp(name, "tophat_x")) { *n = g->tophat_x ? g->ndimx : 0; return g->tophat_x; }
  if (!strcmp(name, "hv_r")) { *n = g->hv_r ? g->p.nrchop * g->ntchopdim : 0; return g->hv_r; }
  if (!strcmp(name, "hv_t")) { *n = g->hv_t ? g->ntchopdim : 0; return g->hv_t; }
  if (!strcmp(name, "hv_x")) { *n = g->hv_x ? g->nxchopdim : 0; return g->hv_x; }
  if (!strcmp(name, "hd_r")) { *n = g->hd_r ? g->p.nrchop * g->ntchopdim : 0; return g->hd_r; }
  if (!strcmp(name, "hd_t")) { *n = g->hd_t ? g->ntchopdim : 0; return g->hd_t; }
  if (!strcmp(name, "hd_x")) { *n = g->hd_x ? g->nxchopdim : 0; return g->hd_x; }
  if (!strcmp(name, "hv_r") && g->hv_r) { *n = g->p.nrchop * g->ntchopdim; return g->hv_r; }
  if (!strcmp(name, "hv_t") && g->hv_t) { *n = g->ntchopdim; return g->hv_t; }
  if (!strcmp(name, "hv_x") && g->hv_x) { *n = g->nxchopdim; return g->hv_x; }
  if (!strcmp(name, "hd_r") && g->hd_r) { *n = g->p.nrchop * g->ntchopdim; return g->hd_r; }
  *n = 0;
  return NULL;
}

/* CHOPSET: mod_scalar3.f90:165-211 */
void og_chopset(og_grid *g, int iof) {
  g->nrchop += iof;
  if (g->nrchop > g->nrchopdim) {
    fprintf(stderr, "og_chopset: NRCHOP too large (%d > %d)\n", g->nrchop, g->nrchopdim);
    abort();
  }
  for (int mm = 0; mm < g->ntchopdim; mm++) g->nrchops[mm] = imax(imin(g->nrchop, g->nrchop - g->m[mm]), 0);
}

int og_nrchop(const og_grid *g) { return g->nrchop; }

/* ===================================================================== */
/* Fields: SALLOC/SFREE/COPY0/CHOPDO (mod_fft.f90:603-769,               */
/* mod_scalar3.f90:213-257)                                               */
/* ===================================================================== */
long og_field_size(const og_grid *g, int space) {
  if (space == OG_FFF) return (long)g->nrchopdim * g->ntchopdim * g->nxchopdim;
  if (space == OG_PFF) return (long)g->ndimr * g->ntchopdim * g->nxchopdim;
  return (long)g->ndimr * g->ndimth * g->ndimx;
}

static long maxsize(const og_grid *g) {
  long a = og_field_size(g, OG_FFF), b = og_field_size(g, OG_PFF), c = og_field_size(g, OG_PPP);
  long s = a > b ? a : b;
  return s > c ? s : c;
}

/* Buffers are recycled through a small per-grid pool (the reference
 * re-allocates on every transform; recycling only removes page-fault cost).
 * A recycled buffer has its FFF-sized head cleared; every caller overwrites
 * or chops the rest, as with a fresh Fortran ALLOCATE. */
#define POOL_MAX 48
static cplx *pool_buf[POOL_MAX];
static long pool_len[POOL_MAX];
static int pool_n = 0;

og_field *og_field_new(og_grid *g, int space) {
  og_field *f = (og_field *)calloc(1, sizeof(og_field));
  long need = maxsize(g);
  f->e = NULL;
  for (int i = pool_n - 1; i >= 0; i--)
    if (pool_len[i] == need) {
      f->e = pool_buf[i];
      pool_buf[i] = pool_buf[pool_n - 1];
      pool_len[i] = pool_len[pool_n - 1];
      pool_n--;
      memset(f->e, 0, sizeof(cplx) * (size_t)og_field_size(g, OG_FFF));
      break;
    }
  if (!f->e) f->e = (cplx *)calloc((size_t)need, sizeof(cplx));
  f->len = need;
  f->space = space;
  f->ln = 0.0;
  return f;
}

void og_field_free(og_field *f) {
  if (!f) return;
  if (pool_n < POOL_MAX) {
    pool_buf[pool_n] = f->e;
    pool_len[pool_n] = f->len;
    pool_n++;
  } else {
    free(f->e);
  }
  free(f);
}

void og_pool_drain(void) {
  for (int i = 0; i < pool_n; i++) free(pool_buf[i]);
  pool_n = 0;
}

/* Re-tag a field the way DEALLOCATE+ALLOCATE does (LN reset; contents are
 * don't-care and are overwritten or chopped by every caller). */
static void realloc_space(const og_grid *g, og_field *f, int space) {
  (void)g;
  f->space = space;
  f->ln = 0.0;
}

void og_chopdo(const og_grid *g, og_field *a) {
  if (a->space == OG_FFF) {
    int lim = imin(g->ntchopdim, g->ntchop);
#pragma omp parallel for
    for (int kk = 0; kk < g->nxchopdim; kk++) {
      for (int mm = 0; mm < lim; mm++)
        for (int n = g->nrchops[mm] + (mutation == 9); n < g->nrchopdim; n++) a->e[FIDX(g, n, mm, kk)] = 0;
      for (int mm = imax(0, g->ntchop); mm < g->ntchopdim; mm++)
        for (int n = 0; n < g->nrchopdim; n++) a->e[FIDX(g, n, mm, kk)] = 0;
    }
  } else if (a->space == OG_PFF) {
#pragma omp parallel for
    for (int kk = 0; kk < g->nxchopdim; kk++)
      for (int mm = 0; mm < g->ntchopdim; mm++)
        for (int r = g->nr; r < g->ndimr; r++) a->e[QIDX(g, r, mm, kk)] = 0;
  } else {
#pragma omp parallel for
    for (int z = 0; z < g->ndimx; z++) {
      if (a->space == OG_PPP && g->ndimth > g->nth)
        for (int r = 0; r < g->ndimr; r++) a->e[PIDX(g, r, g->nth, z)] = 0;
      for (int j = 0; j < g->ndimth; j++)
        for (int r = g->nr; r < g->ndimr; r++) a->e[PIDX(g, r, j, z)] = 0;
    }
  }
}

void og_copy(const og_grid *g, og_field *b, const og_field *a) {
  if (b == a) return;
  if (b->space != a->space) realloc_space(g, b, a->space);
  b->ln = a->ln;
  memcpy(b->e, a->e, sizeof(cplx) * (size_t)og_field_size(g, a->space));
  if (a->space == OG_FFF) og_chopdo(g, b);
}

static void fzero(const og_grid *g, og_field *f) { memset(f->e, 0, sizeof(cplx) * (size_t)og_field_size(g, f->space)); }

/* ===================================================================== */
/* HORFFT: mod_fft.f90:771-926                                           */
/* ===================================================================== */
int og_horfft(const og_grid *g, og_field *a, int is) {
  og_chopdo(g, a);
  int nth = g->nth, n2 = 2 * nth;
  if (is == -1) {
    if (a->space != OG_PPP) return E_TRANSFORM;
    if (nth == 1) { a->space = OG_PFP; return E_OK; }
#pragma omp parallel
    {
      cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n2);
#pragma omp for collapse(2)
      for (int z = 0; z < g->ndimx; z++)
        for (int r = 0; r < g->ndimr; r++) {
          for (int j = 0; j < nth; j++) {
            cplx v = a->e[PIDX(g, r, j, z)];
            buf[2 * j] = creal(v);
            buf[2 * j + 1] = cimag(v);
          }
          fft_c(buf, n2, -1);
          for (int j = 0; j <= nth && j < g->ndimth; j++) a->e[PIDX(g, r, j, z)] = buf[j] / (double)n2;
        }
      free(buf);
    }
#pragma omp parallel for
    for (int z = 0; z < g->ndimx; z++) {
      for (int j = g->ntchop; j < g->ndimth; j++)
        for (int r = 0; r < g->ndimr; r++) a->e[PIDX(g, r, j, z)] = 0;
      for (int r = 0; r < g->ndimr; r++) a->e[PIDX(g, r, 0, z)] = creal(a->e[PIDX(g, r, 0, z)]);
    }
    a->space = OG_PFP;
  } else {
    if (a->space != OG_PFP) return E_TRANSFORM;
    if (nth == 1) { a->space = OG_PPP; return E_OK; }
#pragma omp parallel for
    for (int z = 0; z < g->ndimx; z++) {
      for (int r = 0; r < g->ndimr; r++) a->e[PIDX(g, r, 0, z)] = creal(a->e[PIDX(g, r, 0, z)]);
      for (int j = g->ntchop; j < g->ndimth; j++)
        for (int r = 0; r < g->ndimr; r++) a->e[PIDX(g, r, j, z)] = 0;
    }
#pragma omp parallel
    {
      cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n2);
#pragma omp for collapse(2)
      for (int z = 0; z < g->ndimx; z++)
        for (int r = 0; r < g->ndimr; r++) {
          /* complex-to-real: Hermitian extension, imaginary parts of the
           * zero and Nyquist slots ignored */
          buf[0] = creal(a->e[PIDX(g, r, 0, z)]);
          for (int j = 1; j < nth; j++) {
            cplx v = a->e[PIDX(g, r, j, z)];
            buf[j] = v;
            buf[n2 - j] = conj(v);
          }
          buf[nth] = creal(a->e[PIDX(g, r, nth, z)]);
          fft_c(buf, n2, +1);
          for (int j = 0; j < nth; j++) a->e[PIDX(g, r, j, z)] = creal(buf[2 * j]) + I * creal(buf[2 * j + 1]);
          a->e[PIDX(g, r, nth, z)] = 0;
        }
      free(buf);
    }
    a->space = OG_PPP;
  }
  og_chopdo(g, a);
  return E_OK;
}

/* ===================================================================== */
/* VERFFT: mod_fft.f90:928-1113 (the two MPI exchanges collapse to local  */
/* layout changes on one device)                                          */
/* ===================================================================== */
int og_verfft(const og_grid *g, og_field *a, int is) {
  double ln = a->ln;
  og_chopdo(g, a);
  int nx = g->nx, nxc = g->nxchop, nxd = g->nxchopdim, nxh = g->nxchoph;
  long nout = og_field_size(g, is == -1 ? OG_PFF : OG_PFP);
  cplx *out = (cplx *)calloc((size_t)nout, sizeof(cplx));
  if (is == -1) {
    if (a->space != OG_PFP) { free(out); return E_TRANSFORM; }
    if (nx == 1) { free(out); a->space = OG_PFF; return E_OK; }
#pragma omp parallel
    {
      cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for collapse(2)
      for (int mm = 0; mm < g->ntchopdim; mm++)
        for (int r = 0; r < g->ndimr; r++) {
          for (int z = 0; z < nx; z++) buf[z] = a->e[PIDX(g, r, mm, z)];
          fft_c(buf, nx, -1);
          for (int z = 0; z < nx; z++) buf[z] = buf[z] / (double)nx;
          for (int kk = 0; kk < nxc; kk++) out[QIDX(g, r, mm, kk)] = buf[kk];
          for (int kk = nxc; kk < nxd; kk++) out[QIDX(g, r, mm, kk)] = buf[nxh - 1 + (kk - nxc)];
        }
      free(buf);
    }
    memcpy(a->e, out, sizeof(cplx) * (size_t)nout);
    a->space = OG_PFF;
    a->ln = ln;
  } else {
    if (a->space != OG_PFF) { free(out); return E_TRANSFORM; }
    if (nx == 1) { free(out); a->space = OG_PFP; return E_OK; }
#pragma omp parallel
    {
      cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for collapse(2)
      for (int mm = 0; mm < g->ntchopdim; mm++)
        for (int r = 0; r < g->ndimr; r++) {
          for (int z = 0; z < nx; z++) buf[z] = 0;
          for (int kk = 0; kk < nxc; kk++) buf[kk] = a->e[QIDX(g, r, mm, kk)];
          for (int kk = nxc; kk < nxd; kk++) buf[nxh - 1 + (kk - nxc)] = a->e[QIDX(g, r, mm, kk)];
          fft_c(buf, nx, +1);
          for (int z = 0; z < nx; z++) out[PIDX(g, r, mm, z)] = buf[z];
        }
      free(buf);
    }
    memcpy(a->e, out, sizeof(cplx) * (size_t)nout);
    a->space = OG_PFP;
    a->ln = ln;
  }
  free(out);
  og_chopdo(g, a);
  return E_OK;
}

/* ===================================================================== */
/* RTRAN: mod_fft.f90:1150-1340.  Parity-folded per-m Legendre products.  */
/* ===================================================================== */
int og_rtran(const og_grid *g, og_field *a, int is) {
  int nrh = g->nrh, nr = g->nr, nk = g->nxchopdim;
  if (is == -1) {
    if (a->space != OG_PFF) return E_TRANSFORM;
    cplx *out = (cplx *)calloc((size_t)og_field_size(g, OG_FFF), sizeof(cplx));
#pragma omp parallel
    {
      double *ber = (double *)malloc(sizeof(double) * (size_t)nk * nrh), *bei = (double *)malloc(sizeof(double) * (size_t)nk * nrh);
      double *bor = (double *)malloc(sizeof(double) * (size_t)nk * nrh), *boi = (double *)malloc(sizeof(double) * (size_t)nk * nrh);
      double *sr = (double *)malloc(sizeof(double) * (size_t)nk), *si = (double *)malloc(sizeof(double) * (size_t)nk);
#pragma omp for schedule(dynamic)
      for (int mm = 0; mm < g->ntchopdim; mm++) {
        int nn = g->nrchops[mm];
        if (nn < 1) continue;
        for (int i = 0; i < nrh; i++)
          for (int k = 0; k < nk; k++) {
            cplx p = a->e[QIDX(g, i, mm, k)], q = a->e[QIDX(g, nr - 1 - i, mm, k)];
            cplx e = (p + q) * g->w[i], o = (p - q) * g->w[i];
            ber[k + (size_t)nk * i] = creal(e); bei[k + (size_t)nk * i] = cimag(e);
            bor[k + (size_t)nk * i] = creal(o); boi[k + (size_t)nk * i] = cimag(o);
          }
        for (int n = 0; n < nn; n++) {
          const double *xr = (n % 2 == 0) ? ber : bor, *xi = (n % 2 == 0) ? bei : boi;
          for (int k = 0; k < nk; k++) sr[k] = si[k] = 0.0;
          for (int i0 = 0; i0 < nrh; i0++) {
            int i = sum_rev ? nrh - 1 - i0 : i0;
            double pf = PF(g, i, n, mm);
            const double *ur = xr + (size_t)nk * i, *ui = xi + (size_t)nk * i;
            for (int k = 0; k < nk; k++) { sr[k] += ur[k] * pf; si[k] += ui[k] * pf; }
          }
          for (int k = 0; k < nk; k++) out[FIDX(g, n, mm, k)] = sr[k] + I * si[k];
        }
      }
      free(ber); free(bei); free(bor); free(boi); free(sr); free(si);
    }
    memcpy(a->e, out, sizeof(cplx) * (size_t)og_field_size(g, OG_FFF));
    free(out);
    a->space = OG_FFF;
    a->ln = 0.0;
  } else {
    if (a->space != OG_FFF) return E_TRANSFORM;
    cplx *out = (cplx *)calloc((size_t)og_field_size(g, OG_PFF), sizeof(cplx));
#pragma omp parallel
    {
      double *be_r = (double *)malloc(sizeof(double) * (size_t)nrh), *be_i = (double *)malloc(sizeof(double) * (size_t)nrh);
      double *bo_r = (double *)malloc(sizeof(double) * (size_t)nrh), *bo_i = (double *)malloc(sizeof(double) * (size_t)nrh);
#pragma omp for collapse(2) schedule(dynamic)
      for (int mm = 0; mm < g->ntchopdim; mm++)
        for (int k = 0; k < nk; k++) {
          int nn = g->nrchops[mm];
          for (int i = 0; i < nrh; i++) be_r[i] = be_i[i] = bo_r[i] = bo_i[i] = 0.0;
          for (int n0 = 0; n0 < nn; n0++) {
            int n = sum_rev ? nn - 1 - n0 : n0;
            cplx c = a->e[FIDX(g, n, mm, k)];
            double cr = creal(c), ci = cimag(c);
            const double *pf = &PF(g, 0, n, mm);
            double *tr = (n % 2 == 0) ? be_r : bo_r, *ti = (n % 2 == 0) ? be_i : bo_i;
            for (int i = 0; i < nrh; i++) { tr[i] += pf[i] * cr; ti[i] += pf[i] * ci; }
          }
          for (int i = 0; i < nrh; i++) {
            out[QIDX(g, i, mm, k)] = (be_r[i] + bo_r[i]) + I * (be_i[i] + bo_i[i]);
            out[QIDX(g, nr - 1 - i, mm, k)] = (be_r[i] - bo_r[i]) + I * (be_i[i] - bo_i[i]);
          }
        }
      free(be_r); free(be_i); free(bo_r); free(bo_i);
    }
    for (int mm = imax(0, g->ntchop); mm < g->ntchopdim; mm++)
      for (int k = 0; k < nk; k++)
        for (int r = 0; r < g->ndimr; r++) out[QIDX(g, r, mm, k)] = 0;
    if (a->ln != 0.0)
      for (int i = 0; i < nr; i++) out[QIDX(g, i, 0, 0)] += a->ln * g->lnx[i];
    memcpy(a->e, out, sizeof(cplx) * (size_t)og_field_size(g, OG_PFF));
    free(out);
    a->space = OG_PFF;
    a->ln = 0.0;
  }
  og_chopdo(g, a);
  return E_OK;
}

/* TOFF / TOFP: mod_fft.f90:1342-1398 */
int og_toff(const og_grid *g, og_field *a) {
  int rc = 0;
  if (a->space == OG_PPP) {
    if ((rc = og_horfft(g, a, -1))) return rc;
    if ((rc = og_verfft(g, a, -1))) return rc;
    return og_rtran(g, a, -1);
  } else if (a->space == OG_PFP) {
    if ((rc = og_verfft(g, a, -1))) return rc;
    return og_rtran(g, a, -1);
  } else if (a->space == OG_PFF) {
    return og_rtran(g, a, -1);
  }
  return E_OK;
}

int og_tofp(const og_grid *g, og_field *a) {
  int rc = 0;
  if (a->space == OG_FFF) {
    if ((rc = og_rtran(g, a, 1))) return rc;
    if ((rc = og_verfft(g, a, 1))) return rc;
    return og_horfft(g, a, 1);
  } else if (a->space == OG_PFF) {
    if ((rc = og_verfft(g, a, 1))) return rc;
    return og_horfft(g, a, 1);
  } else if (a->space == OG_PFP) {
    return og_horfft(g, a, 1);
  }
  return E_OK;
}

/* CALCAT1 (value at r = infinity for every k): mod_scalar3.f90:286-310 */
static cplx calcat1(const og_grid *g, const og_field *a, int kk) {
  cplx s = 0;
  for (int n = 0; n < g->nrchop; n++) s += a->e[FIDX(g, n, 0, kk)] * g->at1[n];
  return s;
}

/* ===================================================================== */
/* Spectral operators: mod_legops.f90                                     */
/* ===================================================================== */
/* DELSQH: mod_legops.f90:76-133 */
int og_delsqh(const og_grid *g, const og_field *a, og_field *b) {
  if (a->space != OG_FFF) return E_LEGOP;
  if (b->space != OG_FFF) realloc_space(g, b, OG_FFF);
  b->ln = 0.0;
  og_chopdo(g, b);
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++)
      for (int nn = 0; nn < g->nrchops[mm]; nn++) {
        int n = g->m[mm] + nn;
        b->e[FIDX(g, nn, mm, kk)] = -a->e[FIDX(g, nn, mm, kk)] * ((double)(n * (n + 1)) / g->ell2);
      }
  b->e[FIDX(g, 0, 0, 0)] = a->ln / g->ell2 / NORM(g, 0, 0);
  return E_OK;
}

/* IDELSQH: mod_legops.f90:136-239 */
int og_idelsqh(const og_grid *g, const og_field *b, og_field *a) {
  if (b->space != OG_FFF) return E_LEGOP;
  if (a->space != OG_FFF) realloc_space(g, a, OG_FFF);
  og_chopdo(g, a);
  a->ln = 0.0;
  a->ln = creal(b->e[FIDX(g, 0, 0, 0)]) * NORM(g, 0, 0) * g->ell2;
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++)
      for (int nn = 0; nn < g->nrchops[mm]; nn++) {
        int n = g->m[mm] + nn;
        if (n == 0)
          a->e[FIDX(g, 0, 0, kk)] = 0;
        else
          a->e[FIDX(g, nn, mm, kk)] = -b->e[FIDX(g, nn, mm, kk)] * (g->ell2 / (double)(n * (n + 1)));
      }
  for (int kk = 0; kk < g->nxchopdim; kk++) a->e[FIDX(g, 0, 0, kk)] = a->e[FIDX(g, 0, 0, kk)] - calcat1(g, a, kk) / g->at1[0];
  return E_OK;
}

typedef void (*bandfn)(int, int, const double *, double *);

static void apply_band3(const og_grid *g, const og_field *a, og_field *b, bandfn fn) {
  int ndim = g->nrchopdim + 2;
#pragma omp parallel
  {
    double *band = (double *)malloc(sizeof(double) * 3 * (size_t)ndim);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)ndim);
#pragma omp for
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      int nn = g->nrchops[mm];
      if (nn < 1) continue;
      fn(nn, g->m[mm], &LOGNORM(g, 0, mm), band);
      for (int kk = 0; kk < g->nxchopdim; kk++) {
        banmul(band, nn, 3, 1, &a->e[FIDX(g, 0, mm, kk)], nn, tmp);
        memcpy(&b->e[FIDX(g, 0, mm, kk)], tmp, sizeof(cplx) * (size_t)nn);
      }
    }
    free(band);
    free(tmp);
  }
}

/* MULXM: mod_legops.f90:371-428 */
int og_mulxm(const og_grid *g, const og_field *a, og_field *b) {
  if (a->space != OG_FFF) return E_LEGOP;
  if (b->space != OG_FFF) realloc_space(g, b, OG_FFF);
  b->ln = 0.0;
  og_chopdo(g, b);
  apply_band3(g, a, b, band_xm);
  return E_OK;
}

/* XXDX: mod_legops.f90:680-748 */
int og_xxdx(const og_grid *g, const og_field *a, og_field *b) {
  if (a->space != OG_FFF) return E_LEGOP;
  if (b->space != OG_FFF) realloc_space(g, b, OG_FFF);
  double ln = a->ln;
  b->ln = 0.0;
  og_chopdo(g, b);
  apply_band3(g, a, b, band_xxdx);
  b->e[FIDX(g, 0, 0, 0)] = b->e[FIDX(g, 0, 0, 0)] + ln / NORM(g, 0, 0);
  b->e[FIDX(g, 1, 0, 0)] = b->e[FIDX(g, 1, 0, 0)] + ln / NORM(g, 1, 0);
  return E_OK;
}

/* DEL2: mod_legops.f90:751-820 */
int og_del2(const og_grid *g, const og_field *a, og_field *b) {
  if (a->space != OG_FFF) return E_LEGOP;
  if (b->space != OG_FFF) realloc_space(g, b, OG_FFF);
  if (g->ntchopdim >= 2 && g->nrchops[1] >= 1 && cabs(a->e[FIDX(g, g->nrchops[1] - 1, 1, 0)]) > 1e-15) return E_LEGOP;
  double ln = a->ln;
  b->ln = 0.0;
  og_chopdo(g, b);
  int ndim = g->nrchopdim + 2;
#pragma omp parallel
  {
    double *band = (double *)malloc(sizeof(double) * 5 * (size_t)ndim);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)ndim);
#pragma omp for
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      int nn = g->nrchops[mm];
      if (nn < 1) continue;
      band_del2h(nn, g->m[mm], g->ell, &LOGNORM(g, 0, mm), band);
      for (int kk = 0; kk < g->nxchopdim; kk++) {
        const cplx *x = &a->e[FIDX(g, 0, mm, kk)];
        banmul(band, nn, 5, 2, x, nn, tmp);
        double k2 = pow(AK(g, mm, kk), 2.0);
        for (int i = 0; i < nn; i++) b->e[FIDX(g, i, mm, kk)] = tmp[i] - k2 * x[i];
      }
    }
    free(band);
    free(tmp);
  }
  b->e[FIDX(g, 0, 0, 0)] = b->e[FIDX(g, 0, 0, 0)] + 4.0 / 3 * ln / g->ell2 / NORM(g, 0, 0);
  b->e[FIDX(g, 1, 0, 0)] = b->e[FIDX(g, 1, 0, 0)] - 2.0 * ln / g->ell2 / NORM(g, 1, 0);
  b->e[FIDX(g, 2, 0, 0)] = b->e[FIDX(g, 2, 0, 0)] + 2.0 / 3 * ln / g->ell2 / NORM(g, 2, 0);
  return E_OK;
}

/* The m >= 1 part shared by IDEL2, IDEL2LN and IHELM: penta-diagonal
 * (del^2_h - alp - k^2) factored and solved per (m,k). */
static void solve_penta_mge1(const og_grid *g, og_field *a, double alp) {
  int ndim = g->nrchopdim + 2;
#pragma omp parallel
  {
    double *band = (double *)malloc(sizeof(double) * 5 * (size_t)ndim);
    double *op = (double *)malloc(sizeof(double) * 5 * (size_t)ndim);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)ndim);
#pragma omp for collapse(2) schedule(dynamic)
    for (int mm = 1; mm < g->ntchopdim; mm++)
      for (int kk = 0; kk < g->nxchopdim; kk++) {
        int nn = g->nrchops[mm];
        if (nn < 1) continue;
        band_del2h(nn, g->m[mm], g->ell, &LOGNORM(g, 0, mm), band);
        for (int i = 0; i < nn; i++) band[i + (size_t)nn * 2] = band[i + (size_t)nn * 2] - alp;
        memcpy(op, band, sizeof(double) * 5 * (size_t)nn);
        double k2 = pow(AK(g, mm, kk), 2.0);
        for (int i = 0; i < nn; i++) op[i + (size_t)nn * 2] = band[i + (size_t)nn * 2] - k2;
        lub(op, nn, nn, 5, 2);
        memcpy(x, &a->e[FIDX(g, 0, mm, kk)], sizeof(cplx) * (size_t)nn);
        solveb(op, nn, nn, 5, 2, x);
        memcpy(&a->e[FIDX(g, 0, mm, kk)], x, sizeof(cplx) * (size_t)nn);
      }
    free(band);
    free(op);
    free(x);
  }
}

static void fix_m0_at_infinity(const og_grid *g, og_field *a) {
  for (int kk = 0; kk < g->nxchopdim; kk++) {
    cplx s = 0;
    for (int n = 0; n < g->nrchop; n++) s += g->at1[n] * a->e[FIDX(g, n, 0, kk)];
    a->e[FIDX(g, 0, 0, kk)] = a->e[FIDX(g, 0, 0, kk)] - s / g->at1[0];
  }
}

/* IDEL2: mod_legops.f90:823-1010 */
int og_idel2(const og_grid *g, const og_field *b, og_field *a, int has_ln, double ln) {
  if (b->space != OG_FFF) return E_LEGOP;
  og_copy(g, a, b);
  a->ln = 0.0;
  int nn = g->nrchops[0], np = nn + 1, nm = nn - 1;
  double *op0 = (double *)malloc(sizeof(double) * 5 * (size_t)np);
  band_del2h(np, g->m[0], g->ell, &LOGNORM(g, 0, 0), op0);
#pragma omp parallel
  {
    double *d7 = (double *)malloc(sizeof(double) * 7 * (size_t)np);
    double *sh = (double *)malloc(sizeof(double) * 7 * (size_t)np);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)np);
#pragma omp for
    for (int kk = 0; kk < g->nxchopdim; kk++) {
      memset(d7, 0, sizeof(double) * 7 * (size_t)np);
      for (int i = 0; i < np; i++) {
        d7[i + (size_t)np * 1] = op0[i + (size_t)np * 0];
        d7[i + (size_t)np * 2] = op0[i + (size_t)np * 1];
        d7[i + (size_t)np * 3] = op0[i + (size_t)np * 2] - pow(AK(g, 0, kk), 2.0);
        d7[i + (size_t)np * 4] = op0[i + (size_t)np * 3];
        d7[i + (size_t)np * 5] = op0[i + (size_t)np * 4];
      }
      memcpy(x, &a->e[FIDX(g, 0, 0, kk)], sizeof(cplx) * (size_t)nn);
      if (kk == 0) {
        d7[0 + (size_t)np * 3] = d7[0 + (size_t)np * 3] + 4.0 / 3 / g->ell2;
        d7[1 + (size_t)np * 2] = d7[1 + (size_t)np * 2] - 2 * NORM(g, 0, 0) / NORM(g, 1, 0) / g->ell2;
        d7[2 + (size_t)np * 1] = d7[2 + (size_t)np * 1] + 2.0 / 3 * NORM(g, 0, 0) / NORM(g, 2, 0) / g->ell2;
        memcpy(sh, d7, sizeof(double) * 7 * (size_t)np);
        for (int i = 1; i < np; i++)
          for (int c = 0; c < 5; c++) d7[i + (size_t)np * c] = sh[i - 1 + (size_t)np * (c + 1)];
        for (int i = 0; i < np; i++) { d7[i + (size_t)np * 5] = 0; d7[i + (size_t)np * 6] = 0; }
        for (int c = 0; c < 7; c++) d7[0 + (size_t)np * c] = 0;
        d7[0 + (size_t)np * 3] = 1;
        for (int i = nm; i >= 1; i--) x[i] = x[i - 1];
        x[0] = has_ln ? ln / NORM(g, 0, 0) : 0.0;
      }
      lub(d7, np, nn, 7, 3);
      solveb(d7, np, nn, 7, 3, x);
      memcpy(&a->e[FIDX(g, 0, 0, kk)], x, sizeof(cplx) * (size_t)nn);
    }
    free(d7);
    free(sh);
    free(x);
  }
  free(op0);
  solve_penta_mge1(g, a, 0.0);
  if (has_ln) a->ln = creal(a->e[FIDX(g, 0, 0, 0)]) * NORM(g, 0, 0);
  og_chopdo(g, a);
  fix_m0_at_infinity(g, a);
  return E_OK;
}

/* IDEL2LN: mod_legops.f90:1013-1176 */
int og_idel2ln(const og_grid *g, const og_field *b, og_field *a) {
  if (b->space != OG_FFF) return E_LEGOP;
  og_copy(g, a, b);
  a->ln = 0.0;
  int nn = g->nrchops[0], np = nn + 1;
  double *op0 = (double *)malloc(sizeof(double) * 5 * (size_t)np);
  band_del2h(np, g->m[0], g->ell, &LOGNORM(g, 0, 0), op0);
  double ln00 = 0.0;
#pragma omp parallel
  {
    double *d5 = (double *)malloc(sizeof(double) * 5 * (size_t)np);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)np);
#pragma omp for
    for (int kk = 0; kk < g->nxchopdim; kk++) {
      memcpy(d5, op0, sizeof(double) * 5 * (size_t)np);
      for (int i = 0; i < np; i++) d5[i + (size_t)np * 2] = op0[i + (size_t)np * 2] - pow(AK(g, 0, kk), 2.0);
      if (kk == 0) {
        d5[0 + (size_t)np * 2] = d5[0 + (size_t)np * 2] + 4.0 / 3 / g->ell2;
        d5[1 + (size_t)np * 1] = d5[1 + (size_t)np * 1] - 2 * NORM(g, 0, 0) / NORM(g, 1, 0) / g->ell2;
        d5[2 + (size_t)np * 0] = d5[2 + (size_t)np * 0] + 2.0 / 3 * NORM(g, 0, 0) / NORM(g, 2, 0) / g->ell2;
      }
      memcpy(x, &a->e[FIDX(g, 0, 0, kk)], sizeof(cplx) * (size_t)nn);
      lub(d5, np, nn, 5, 2);
      solveb(d5, np, nn, 5, 2, x);
      memcpy(&a->e[FIDX(g, 0, 0, kk)], x, sizeof(cplx) * (size_t)nn);
    }
    free(d5);
    free(x);
  }
  free(op0);
  ln00 = creal(a->e[FIDX(g, 0, 0, 0)]) * NORM(g, 0, 0);
  a->e[FIDX(g, 0, 0, 0)] = 0;
  solve_penta_mge1(g, a, 0.0);
  a->ln = ln00;
  og_chopdo(g, a);
  fix_m0_at_infinity(g, a);
  return E_OK;
}

/* IHELM: mod_legops.f90:1204-1354 */
int og_ihelm(const og_grid *g, const og_field *b, og_field *a, double alp) {
  if (b->space != OG_FFF) return E_LEGOP;
  if (alp == 0.0) return E_LEGOP;
  double bln = b->ln;
  og_copy(g, a, b);
  a->ln = -bln / alp;
  a->e[FIDX(g, 0, 0, 0)] = a->e[FIDX(g, 0, 0, 0)] - a->ln * (4.0 / 3 / g->ell2) / NORM(g, 0, 0);
  a->e[FIDX(g, 1, 0, 0)] = a->e[FIDX(g, 1, 0, 0)] - a->ln * (-2 / NORM(g, 1, 0) / g->ell2);
  a->e[FIDX(g, 2, 0, 0)] = a->e[FIDX(g, 2, 0, 0)] - a->ln * (+2.0 / 3 / NORM(g, 2, 0) / g->ell2);
  int nn = g->nrchops[0], np = nn + 1, nm = nn - 1;
  double *xm = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  band_xm(np, g->m[0], &LOGNORM(g, 0, 0), xm);
  double *h0 = (double *)malloc(sizeof(double) * 5 * (size_t)np);
  band_del2h(np, g->m[0], g->ell, &LOGNORM(g, 0, 0), h0);
  for (int i = 0; i < np; i++) h0[i + (size_t)np * 2] = h0[i + (size_t)np * 2] - alp;
#pragma omp parallel
  {
    double *h5 = (double *)malloc(sizeof(double) * 5 * (size_t)np);
    double *h7 = (double *)malloc(sizeof(double) * 7 * (size_t)np);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)np), *y = (cplx *)malloc(sizeof(cplx) * (size_t)np);
#pragma omp for
    for (int kk = 0; kk < g->nxchopdim; kk++) {
      memcpy(h5, h0, sizeof(double) * 5 * (size_t)np);
      for (int i = 0; i < np; i++) h5[i + (size_t)np * 2] = h0[i + (size_t)np * 2] - pow(AK(g, 0, kk), 2.0);
      banbanmul(h5, 5, 2, xm, 3, 1, np, h7);
      lub(h7, np, nm, 7, 3);
      memcpy(x, &a->e[FIDX(g, 0, 0, kk)], sizeof(cplx) * (size_t)nm);
      solveb(h7, np, nm, 7, 3, x);
      x[nn - 1] = 0;
      banmul(xm, np, 3, 1, x, nn, y);
      memcpy(&a->e[FIDX(g, 0, 0, kk)], y, sizeof(cplx) * (size_t)nn);
    }
    free(h5);
    free(h7);
    free(x);
    free(y);
  }
  free(xm);
  free(h0);
  solve_penta_mge1(g, a, alp);
  og_chopdo(g, a);
  return E_OK;
}

/* BANMULR: mod_bandmat.f90 (band x real vector), as banmul. */
static void banmul_r(const double *a, int lda, int nb, int nc, const double *x, int n, double *out) {
  for (int c = 0; c < n; c++) out[c] = a[c + (size_t)lda * nc] * x[c];
  for (int bi = 0; bi < nb; bi++) {
    if (bi == nc) continue;
    int is = bi - nc;
    int c0 = imax(0, -is), c1 = imin(n - 1, n - 1 - is);
    for (int c = c0; c <= c1; c++) out[c] += a[c + (size_t)lda * bi] * x[c + is];
  }
}

/* The (2P+1)-band matrix of (del^P + nu del^2 - alp) for mode mm and axial
 * wavenumber squared k2 (mod_legops.f90:1518-1572): h[np x (2p+1)], rows
 * < nn modified by nu and alp.  d[j] holds the band of del_h^(2j) (width
 * 4j+1, centre 2j) over np rows; the binomial terms are added in the
 * reference's order. */
static void helmp_matrix(int p, int np, int nn, double *const *d, double k2, double nu, double alp, double *h) {
  const int q = p / 2, w = 2 * p + 1;
  static const double binom[8][8] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1}, {1, 5, 10, 10, 5, 1},
                                     {1, 6, 15, 20, 15, 6, 1}, {1, 7, 21, 35, 35, 21, 7, 1}};
  memset(h, 0, sizeof(double) * (size_t)np * w);
  for (int c = 0; c < w; c++)
    for (int r = 0; r < np; r++) h[r + (size_t)np * c] = d[q][r + (size_t)np * c];
  /* - C(q,1) K2 del^(P-2) + C(q,2) K2^2 del^(P-4) - ... */
  for (int t = 1; t < q; t++) {
    const int j = q - t, wj = 4 * j + 1, off = 2 * (q - j);
    double kp = k2;
    for (int e = 1; e < t; e++) kp = kp * k2;
    double cf = (t & 1) ? -(binom[q][t] * kp) : binom[q][t] * kp;
    if (mutation == 20) cf = -cf;
    for (int c = 0; c < wj; c++)
      for (int r = 0; r < np; r++) h[r + (size_t)np * (off + c)] = h[r + (size_t)np * (off + c)] + cf * d[j][r + (size_t)np * c];
  }
  {
    double kp = k2;
    for (int e = 1; e < q; e++) kp = kp * k2;
    for (int r = 0; r < np; r++) h[r + (size_t)np * p] = (q & 1) ? h[r + (size_t)np * p] - kp : h[r + (size_t)np * p] + kp;
  }
  for (int c = 0; c < 5; c++) {
    const int cc = p - 2 + c;
    for (int r = 0; r < nn; r++) {
      double dv = d[1][r + (size_t)np * c];
      if (c == 2) dv = dv - k2;
      h[r + (size_t)np * cc] = h[r + (size_t)np * cc] + (mutation == 21 ? 0.0 : nu) * dv;
    }
  }
  for (int r = 0; r < nn; r++) h[r + (size_t)np * p] = h[r + (size_t)np * p] - alp;
}

/* IHELMP: mod_legops.f90:1409-1592, A = (del^P + nu del^2 - alp)^-1 B for
 * even P in 4..14: the log term's contribution at m = 0, k = 0, then one
 * (2P+1)-band LU solve per (m, k) over NRCHOPS(m) rows (the band powers
 * formed over NRCHOPS(m)+P rows), the value at infinity of m = 0, CHOPDO. */
int og_ihelmp(const og_grid *g, int p, const og_field *b, og_field *a, double alp, double nu) {
  if (b->space != OG_FFF) return E_LEGOP;
  if (g->m[0] != 0 || alp == 0.0) return E_LEGOP;
  if (p < 4 || p > 14 || (p & 1)) return E_LEGOP;
  const double bln = b->ln;
  og_copy(g, a, b);
  a->ln = -bln / alp;
  {
    const int nr = g->nrchop;
    double *al2 = (double *)calloc((size_t)nr + 3, sizeof(double)), *al = (double *)calloc((size_t)nr + 3, sizeof(double));
    double *t = (double *)calloc((size_t)nr + 3, sizeof(double)), *h = (double *)malloc(sizeof(double) * 5 * (size_t)nr);
    al2[0] = a->ln * (4.0 / 3 / g->ell2) / NORM(g, 0, 0);
    al2[1] = a->ln * (-2 / NORM(g, 1, 0) / g->ell2);
    al2[2] = a->ln * (+2.0 / 3 / NORM(g, 2, 0) / g->ell2);
    memcpy(al, al2, sizeof(double) * (size_t)nr);
    band_del2h(nr, g->m[0], g->ell, &LOGNORM(g, 0, 0), h);
    for (int i = 4; i <= p; i += 2) {
      banmul_r(h, nr, 5, 2, al, nr, t);
      memcpy(al, t, sizeof(double) * (size_t)nr);
    }
    for (int n = 0; n < nr; n++) a->e[FIDX(g, n, 0, 0)] = a->e[FIDX(g, n, 0, 0)] - al[n] - (mutation == 22 ? 0.0 : nu) * al2[n];
    free(al2);
    free(al);
    free(t);
    free(h);
  }
  const int w = 2 * p + 1, q = p / 2;
#pragma omp parallel for schedule(dynamic)
  for (int mm = 0; mm < g->ntchopdim; mm++) {
    const int nn = g->nrchops[mm], np = nn + p;
    double *d[8] = {0};
    for (int j = 1; j <= q; j++) d[j] = (double *)malloc(sizeof(double) * (size_t)np * (4 * j + 1));
    band_del2h(np, g->m[mm], g->ell, &LOGNORM(g, 0, mm), d[1]);
    for (int j = 2; j <= q; j++) banbanmul(d[j - 1], 4 * j - 3, 2 * j - 2, d[1], 5, 2, np, d[j]);
    double *h = (double *)malloc(sizeof(double) * (size_t)np * w);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)(nn > 0 ? nn : 1));
    for (int kk = 0; kk < g->nxchopdim; kk++) {
      const double k2 = pow(AK(g, mm, kk), 2.0);
      helmp_matrix(p, np, nn, d, k2, nu, alp, h);
      if (nn < 1) continue;
      lub(h, np, nn, w, p);
      for (int n = 0; n < nn; n++) x[n] = a->e[FIDX(g, n, mm, kk)];
      solveb(h, np, nn, w, p, x);
      for (int n = 0; n < nn; n++) a->e[FIDX(g, n, mm, kk)] = x[n];
    }
    for (int j = 1; j <= q; j++) free(d[j]);
    free(h);
    free(x);
  }
  /* zero at infinity: A(1,1,:) -= CALCAT1(A) / PF(1,1,1) */
  for (int kk = 0; kk < g->nxchopdim; kk++)
    a->e[FIDX(g, 0, 0, kk)] = a->e[FIDX(g, 0, 0, kk)] - calcat1(g, a, kk) / PF(g, 0, 0, 0);
  og_chopdo(g, a);
  return E_OK;
}

/* HELMP: mod_legops.f90:1357-1406, B = (del^P + nu del^2 + alp) A (the
 * forward operator, for checks of IHELMP): repeated DEL2. */
int og_helmp(const og_grid *g, int p, const og_field *a, og_field *b, double alp, double nu) {
  if (a->space != OG_FFF) return E_LEGOP;
  og_field *w = og_field_new((og_grid *)g, OG_FFF), *d2 = og_field_new((og_grid *)g, OG_FFF);
  int rc = og_del2(g, a, d2);
  if (!rc) og_copy(g, b, d2);
  for (int i = 4; !rc && i <= p; i += 2) {
    rc = og_del2(g, b, w);
    if (!rc) og_copy(g, b, w);
  }
  if (!rc) {
    for (size_t i = 0; i < (size_t)g->nrchopdim * g->ntchopdim * g->nxchopdim; i++) {
      b->e[i] = b->e[i] - alp * a->e[i];
      b->e[i] = b->e[i] + nu * d2->e[i];
    }
    b->ln = -alp * a->ln;
  }
  og_field_free(w);
  og_field_free(d2);
  return rc;
}

/* GRAD: mod_legops.f90:1727-1786 */
int og_grad(const og_grid *g, const og_field *a, og_field *rr, og_field *rp, og_field *z) {
  if (rr->space != OG_FFF) { realloc_space(g, rr, OG_FFF); realloc_space(g, rp, OG_FFF); realloc_space(g, z, OG_FFF); }
  int rc = og_xxdx(g, a, rr);
  if (rc) return rc;
  if (rp->space != OG_FFF) realloc_space(g, rp, OG_FFF);
  if (z->space != OG_FFF) realloc_space(g, z, OG_FFF);
  rp->ln = 0.0; og_chopdo(g, rp);
  z->ln = 0.0; og_chopdo(g, z);
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      double mv = g->m[mm], kv = AK(g, mm, kk);
      for (int n = 0; n < g->nrchops[mm]; n++) {
        cplx v = a->e[FIDX(g, n, mm, kk)];
        rp->e[FIDX(g, n, mm, kk)] = (mutation == 7 ? -I * mv : I * mv) * v;
        z->e[FIDX(g, n, mm, kk)] = (mutation == 10 ? -I * kv : I * kv) * v;
      }
    }
  return E_OK;
}

/* PC2VEL: mod_legops.f90:1789-1853 */
int og_pc2vel(const og_grid *g, const og_field *psi, const og_field *chi, og_field *rur, og_field *rup, og_field *uz) {
  int rc;
  if ((rc = og_xxdx(g, chi, rur))) return rc;
  if ((rc = og_xxdx(g, psi, rup))) return rc;
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      double mv = g->m[mm], kv = AK(g, mm, kk);
      for (int n = 0; n < g->nrchops[mm]; n++) {
        size_t id = FIDX(g, n, mm, kk);
        rur->e[id] = (I * mv) * psi->e[id] + (I * kv) * rur->e[id];
        rup->e[id] = -rup->e[id] - (mv * kv) * chi->e[id];
      }
    }
  if ((rc = og_delsqh(g, chi, uz))) return rc;
  og_field *w = og_field_new((og_grid *)g, OG_FFF);
  og_mulxm(g, uz, w);
  og_mulxm(g, w, uz);
  og_field_free(w);
  long n = og_field_size(g, OG_FFF);
  for (long i = 0; i < n; i++) uz->e[i] = -uz->e[i];
  return E_OK;
}

/* PC2VOR: mod_legops.f90:1856-1919 */
int og_pc2vor(const og_grid *g, const og_field *psi, const og_field *chi, og_field *ror, og_field *rop, og_field *oz) {
  int rc;
  if ((rc = og_del2(g, chi, oz))) return rc;
  if ((rc = og_xxdx(g, psi, ror))) return rc;
  if ((rc = og_xxdx(g, oz, rop))) return rc;
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      double mv = g->m[mm], kv = AK(g, mm, kk);
      for (int n = 0; n < g->nrchops[mm]; n++) {
        size_t id = FIDX(g, n, mm, kk);
        ror->e[id] = (mutation == 6 ? (I * mv) : -(I * mv)) * oz->e[id] + (I * kv) * ror->e[id];
        rop->e[id] = rop->e[id] - (kv * mv) * psi->e[id];
      }
    }
  if ((rc = og_delsqh(g, psi, oz))) return rc;
  og_field *w = og_field_new((og_grid *)g, OG_FFF);
  og_mulxm(g, oz, w);
  og_mulxm(g, w, oz);
  og_field_free(w);
  long n = og_field_size(g, OG_FFF);
  for (long i = 0; i < n; i++) oz->e[i] = -oz->e[i];
  return E_OK;
}

/* VPRODSUB cross product: mod_legops.f90:2106-2130 (free stream, strain and
 * the velocity monitor are zero/off on this path). */
int og_vprodsub(const og_grid *g, og_field *rur, og_field *rup, og_field *uz, og_field *ror, og_field *rop, og_field *oz) {
  (void)rur;
#pragma omp parallel for collapse(2)
  for (int z = 0; z < g->ndimx; z++)
    for (int j = 0; j < g->nth; j++)
      for (int r = 0; r < g->nr; r++) {
        size_t id = PIDX(g, r, j, z);
        double a1 = creal(rur->e[id]), a2 = creal(rup->e[id]), a3 = creal(uz->e[id]);
        double c1 = cimag(rur->e[id]), c2 = cimag(rup->e[id]), c3 = cimag(uz->e[id]);
        double b1 = creal(ror->e[id]), b2 = creal(rop->e[id]), b3 = creal(oz->e[id]);
        double d1 = cimag(ror->e[id]), d2 = cimag(rop->e[id]), d3 = cimag(oz->e[id]);
        double rr = g->r[r] * g->r[r];
        ror->e[id] = (a2 * b3 - a3 * b2) + I * (c2 * d3 - c3 * d2);
        rop->e[id] = (a3 * b1 - a1 * b3) + I * (c3 * d1 - c1 * d3);
        oz->e[id] = (a1 * b2 - a2 * b1) / rr + I * ((c1 * d2 - c2 * d1) / rr);
      }
  return E_OK;
}

/* VDOTPSUB: mod_legops.f90:2136-2232 */
int og_vdotpsub(const og_grid *g, const og_field *rar, const og_field *rap, const og_field *az, const og_field *rbr,
                const og_field *rbp, const og_field *bz, og_field *out) {
  out->space = OG_PPP;
  fzero(g, out);
#pragma omp parallel for collapse(2)
  for (int z = 0; z < g->ndimx; z++)
    for (int j = 0; j < g->nth; j++)
      for (int r = 0; r < g->nr; r++) {
        size_t id = PIDX(g, r, j, z);
        double rr = g->r[r] * g->r[r];
        double a1 = creal(rar->e[id]), a2 = creal(rap->e[id]), a3 = creal(az->e[id]);
        double c1 = cimag(rar->e[id]), c2 = cimag(rap->e[id]), c3 = cimag(az->e[id]);
        double b1 = creal(rbr->e[id]), b2 = creal(rbp->e[id]), b3 = creal(bz->e[id]);
        double d1 = cimag(rbr->e[id]), d2 = cimag(rbp->e[id]), d3 = cimag(bz->e[id]);
        out->e[id] = ((a1 * b1 + a2 * b2) / rr + a3 * b3) + I * ((c1 * d1 + c2 * d2) / rr + c3 * d3);
      }
  return E_OK;
}

/* VEL2PLN: mod_legops.f90:2277-2339 */
int og_vel2pln(const og_grid *g, og_field *rup, og_field *psi, int has_ln, double ln) {
  if (rup->space != OG_PFF) return E_LEGOP;
  if (psi->space != OG_FFF) realloc_space(g, psi, OG_FFF);
  psi->ln = 0.0;
  int nrh = g->nrh, nr = g->nr, nn = g->nrchops[0];
  cplx *be = (cplx *)malloc(sizeof(cplx) * (size_t)nrh), *bo = (cplx *)malloc(sizeof(cplx) * (size_t)nrh);
  for (int i = 0; i < nrh; i++) {
    be[i] = (rup->e[QIDX(g, i, 0, 0)] + rup->e[QIDX(g, nr - 1 - i, 0, 0)]) * g->w[i];
    bo[i] = (rup->e[QIDX(g, i, 0, 0)] - rup->e[QIDX(g, nr - 1 - i, 0, 0)]) * g->w[i];
  }
  double s = 0.0;
  for (int n = 0; n < nn; n++) {
    const cplx *src = (n % 2 == 0) ? be : bo;
    double cr = 0.0;
    for (int i = 0; i < nrh; i++) cr += creal(src[i]) * PF(g, i, n, 0);
    s += cr * exp(LOGNORM(g, n, 0));
  }
  psi->ln = -s / 2.0;
  for (int i = 0; i < nrh; i++) {
    double e = -PF(g, i, 0, 0) * psi->ln / NORM(g, 0, 0);
    double o = -PF(g, i, 1, 0) * psi->ln / NORM(g, 1, 0);
    rup->e[QIDX(g, i, 0, 0)] = rup->e[QIDX(g, i, 0, 0)] - (e + o);
    rup->e[QIDX(g, nr - 1 - i, 0, 0)] = rup->e[QIDX(g, nr - 1 - i, 0, 0)] - (e - o);
  }
  free(be);
  free(bo);
  if (has_ln) psi->ln = ln;
  return E_OK;
}

/* PROJECT + EOMUL/OEMUL: mod_legops.f90:2397-2648 */
int og_project(const og_grid *g, og_field *rur, og_field *rup, og_field *uz, og_field *psi, og_field *chi, int has_ln,
               double ln) {
  if (!(rur->space == rup->space && rup->space == uz->space)) return E_TRANSFORM;
  int rc = 0;
  if (rur->space == OG_PPP) {
    og_field *fs[3] = {rur, rup, uz};
    for (int i = 0; i < 3; i++) {
      if ((rc = og_horfft(g, fs[i], -1))) return rc;
      if ((rc = og_verfft(g, fs[i], -1))) return rc;
    }
  } else if (rur->space == OG_FFF) {
    if ((rc = og_rtran(g, rur, 1)) || (rc = og_rtran(g, rup, 1)) || (rc = og_rtran(g, uz, 1))) return rc;
  } else if (rur->space != OG_PFF) {
    return E_TRANSFORM;
  }
  if (psi->space != OG_FFF) realloc_space(g, psi, OG_FFF);
  if (chi->space != OG_FFF) realloc_space(g, chi, OG_FFF);
  if (has_ln) {
    psi->ln = ln;
    if ((rc = og_vel2pln(g, rup, psi, 1, ln))) return rc;
  }
  int nrh = g->nrh, nr = g->nr, nk = g->nxchopdim;
  double *fff = (double *)malloc(sizeof(double) * (size_t)nrh);
  for (int i = 0; i < nrh; i++) fff[i] = (1 - g->x[i] * g->x[i]) / g->w[i];
#pragma omp parallel
  {
    int nmax = g->nrchopdim + 1;
    double *v = (double *)malloc(sizeof(double) * (size_t)nmax * nrh);
    double *d = (double *)malloc(sizeof(double) * (size_t)nmax * nrh);
    double *t = (double *)malloc(sizeof(double) * (size_t)nmax * nrh);
    double *pfd = (double *)malloc(sizeof(double) * (size_t)nmax * nrh);
    double *xb = (double *)malloc(sizeof(double) * 3 * (size_t)(nmax + 1));
    /* fold buffers: [field][even/odd][re/im][k][i] */
    double *fb = (double *)malloc(sizeof(double) * 12 * (size_t)nk * nrh);
    cplx *w1 = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * nk), *w2 = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * nk);
    cplx *w3 = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * nk), *w4 = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * nk);
    cplx *w5 = (cplx *)malloc(sizeof(cplx) * (size_t)nmax * nk);
#pragma omp for schedule(dynamic)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      int mv = g->m[mm], nn = g->nrchops[mm];
      if (nn < 1) continue;
      /* PFD = PF(:, :nn+1) * XXDX band (vector-times-band, VR_BAN_MUL order) */
      band_xxdx(nn + 1, mv, &LOGNORM(g, 0, mm), xb);
      int nb = nn + 1;
      for (int i = 0; i < nrh; i++) {
        for (int c = 0; c < nb; c++) pfd[i + (size_t)nrh * c] = 0.0;
        for (int bi = 0; bi < 3; bi++) {
          int is = bi - 1;
          int c0 = imax(0, -is), c1 = imin(nb - 1, nb - 1 - is);
          for (int c = c0; c <= c1; c++) pfd[i + (size_t)nrh * (c + is)] += xb[c + (size_t)nb * bi] * PF(g, i, c, mm);
        }
      }
      for (int ii = 0; ii < nn; ii++) {
        int n = imax(1, mv + ii);
        for (int j = 0; j < nrh; j++) {
          double den = (double)(n * (n + 1)) * fff[j];
          v[ii + (size_t)nmax * j] = PF(g, j, ii, mm) / den;
          d[ii + (size_t)nmax * j] = pfd[j + (size_t)nrh * ii] / den;
          t[ii + (size_t)nmax * j] = PF(g, j, ii, mm) * g->w[j];
        }
      }
      const og_field *src[3] = {rur, rup, uz};
      for (int f = 0; f < 3; f++)
        for (int j = 0; j < nrh; j++)
          for (int k = 0; k < nk; k++) {
            cplx p = src[f]->e[QIDX(g, j, mm, k)], q = src[f]->e[QIDX(g, nr - 1 - j, mm, k)];
            cplx e = p + q, o = p - q;
            fb[k + (size_t)nk * (j + (size_t)nrh * (0 + 4 * f))] = creal(e);
            fb[k + (size_t)nk * (j + (size_t)nrh * (1 + 4 * f))] = cimag(e);
            fb[k + (size_t)nk * (j + (size_t)nrh * (2 + 4 * f))] = creal(o);
            fb[k + (size_t)nk * (j + (size_t)nrh * (3 + 4 * f))] = cimag(o);
          }
      /* C(n,k) = sum_j A(n,j) * fold(j,k); EO: even n <- even fold, odd n <- odd fold; OE: swapped */
#define MULFOLD(A, f, eo, out)                                                                      \
  for (int ii = 0; ii < nn; ii++) {                                                                 \
    int par = ((ii % 2 == 0) == (eo)) ? 0 : 2;                                                      \
    const double *br = fb + (size_t)nk * nrh * (par + 4 * (f));                                     \
    const double *bi_ = fb + (size_t)nk * nrh * (par + 1 + 4 * (f));                                \
    for (int k = 0; k < nk; k++) {                                                                  \
      double sr = 0, si = 0;                                                                        \
      for (int j0 = 0; j0 < nrh; j0++) {                                                            \
        int j = sum_rev ? nrh - 1 - j0 : j0;                                                        \
        double av = (A)[ii + (size_t)nmax * j];                                                     \
        sr += av * br[k + (size_t)nk * j];                                                          \
        si += av * bi_[k + (size_t)nk * j];                                                         \
      }                                                                                             \
      (out)[ii + (size_t)nmax * k] = sr + I * si;                                                   \
    }                                                                                               \
  }
      for (int q = 0; q < nmax * nk; q++) w1[q] = w2[q] = w3[q] = w4[q] = w5[q] = 0;
      if (mv != 0) { MULFOLD(v, 0, 1, w1); }
      MULFOLD(d, 1, 0, w2);
      MULFOLD(d, 0, 0, w3);
      if (mv != 0) { MULFOLD(v, 1, 1, w4); } else { memcpy(w4, w2, sizeof(cplx) * (size_t)nmax * nk); }
      MULFOLD(t, 2, 1, w5);
#undef MULFOLD
      for (int k = 0; k < nk; k++) {
        double kv = AK(g, mm, k);
        for (int ii = 0; ii < nn; ii++) {
          psi->e[FIDX(g, ii, mm, k)] = -(I * (double)mv) * w1[ii + (size_t)nmax * k] - w2[ii + (size_t)nmax * k];
          chi->e[FIDX(g, ii, mm, k)] =
              (I * kv) * w3[ii + (size_t)nmax * k] + (mv * kv) * w4[ii + (size_t)nmax * k] - w5[ii + (size_t)nmax * k];
        }
      }
    }
    free(v); free(d); free(t); free(pfd); free(xb); free(fb);
    free(w1); free(w2); free(w3); free(w4); free(w5);
  }
  free(fff);
  og_chopdo(g, psi);
  og_chopdo(g, chi);
  for (int kk = 0; kk < g->nxchopdim; kk++) psi->e[FIDX(g, 0, 0, kk)] = psi->e[FIDX(g, 0, 0, kk)] - calcat1(g, psi, kk) / g->at1[0];
  return E_OK;
}

/* ===================================================================== */
/* Boussinesq terms: mod_boussinesq.f90                                   */
/* ===================================================================== */
static void scale_add(const og_grid *g, og_field *y, const og_field *x, double a, int space) {
  long n = og_field_size(g, space);
#pragma omp parallel for
  for (long i = 0; i < n; i++) y->e[i] = y->e[i] + a * x->e[i];
}

/* CALC_BOUSSI_FORCE: mod_boussinesq.f90:317-410 */
int og_force(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b, og_field *psin, og_field *chin,
             og_field *bn) {
  int rc = 0;
  og_field *rur = og_field_new(g, OG_FFF), *rup = og_field_new(g, OG_FFF), *uz = og_field_new(g, OG_FFF);
  og_field *ror = og_field_new(g, OG_FFF), *rop = og_field_new(g, OG_FFF), *oz = og_field_new(g, OG_FFF);
  og_field *rbr = og_field_new(g, OG_FFF), *rbp = og_field_new(g, OG_FFF), *bz = og_field_new(g, OG_FFF);
  og_field *w = og_field_new(g, OG_FFF);
  og_chopset(g, 3);
  if ((rc = og_grad(g, b, rbr, rbp, bz))) goto done;
  if ((rc = og_pc2vel(g, psi, chi, rur, rup, uz))) goto done;
  if ((rc = og_pc2vor(g, psi, chi, ror, rop, oz))) goto done;
  {
    og_field *fs[9] = {rur, rup, uz, ror, rop, oz, rbr, rbp, bz};
    for (int i = 0; i < 9; i++)
      if ((rc = og_tofp(g, fs[i]))) goto done;
  }
  og_vprodsub(g, rur, rup, uz, ror, rop, oz);
  og_vdotpsub(g, rur, rup, uz, rbr, rbp, bz, bn);
  {
    long n = og_field_size(g, OG_PPP);
    for (long i = 0; i < n; i++) bn->e[i] = -bn->e[i];
  }
  bn->ln = 0.0;
  og_chopset(g, -3);
  og_copy(g, w, b);
  if ((rc = og_tofp(g, w))) goto done;
  {
    long n = og_field_size(g, OG_PPP);
    double o2 = 2 * g->p.omega, n2 = pow(g->p.bv0, 2.0);
#pragma omp parallel for
    for (long i = 0; i < n; i++) {
      ror->e[i] = ror->e[i] + o2 * rup->e[i];
      rop->e[i] = mutation == 2 ? rop->e[i] + o2 * rur->e[i] : rop->e[i] - o2 * rur->e[i];
      oz->e[i] = mutation == 1 ? oz->e[i] + w->e[i] : oz->e[i] - w->e[i];
      bn->e[i] = mutation == 3 ? bn->e[i] - n2 * uz->e[i] : bn->e[i] + n2 * uz->e[i];
    }
  }
  realloc_space(g, w, OG_FFF);
  if ((rc = og_project(g, ror, rop, oz, psin, w, 1, 0.0))) goto done;
  if ((rc = og_idel2ln(g, w, chin))) goto done;
  rc = og_toff(g, bn);
done:
  og_field_free(rur); og_field_free(rup); og_field_free(uz);
  og_field_free(ror); og_field_free(rop); og_field_free(oz);
  og_field_free(rbr); og_field_free(rbp); og_field_free(bz);
  og_field_free(w);
  return rc;
}

/* CALC_BOUSSI_NONLIN: mod_boussinesq.f90:235-314.  On exit ror/rop/oz hold
 * (r N_r, r N_theta, N_z) in FFF and bn holds -(u.grad b) in FFF. */
int og_nonlin(og_grid *g, const og_field *b, og_field *rur, og_field *rup, og_field *uz, og_field *bn, og_field *ror,
              og_field *rop, og_field *oz) {
  int rc = 0;
  if (b->space != OG_FFF) return E_LEGOP;
  og_field *rbr = og_field_new(g, OG_FFF), *rbp = og_field_new(g, OG_FFF), *bz = og_field_new(g, OG_FFF);
  og_field *psi = og_field_new(g, OG_FFF), *chi = og_field_new(g, OG_FFF), *w = og_field_new(g, OG_FFF);
  og_chopset(g, 1);
  rc = og_grad(g, b, rbr, rbp, bz);
  og_chopset(g, -1);
  if (rc) goto done;
  {
    og_field *fs[9] = {rur, rup, uz, ror, rop, oz, rbr, rbp, bz};
    for (int i = 0; i < 9; i++)
      if ((rc = og_tofp(g, fs[i]))) goto done;
  }
  og_vprodsub(g, rur, rup, uz, ror, rop, oz);
  og_vdotpsub(g, rur, rup, uz, rbr, rbp, bz, bn);
  {
    long n = og_field_size(g, OG_PPP);
    for (long i = 0; i < n; i++) bn->e[i] = -bn->e[i];
  }
  bn->ln = 0.0;
  if ((rc = og_project(g, ror, rop, oz, psi, w, 1, 0.0))) goto done;
  if ((rc = og_idel2(g, w, chi, 0, 0.0))) goto done;
  og_chopset(g, 3);
  realloc_space(g, ror, OG_FFF); realloc_space(g, rop, OG_FFF); realloc_space(g, oz, OG_FFF);
  rc = og_pc2vel(g, psi, chi, ror, rop, oz);
  og_chopset(g, -3);
  if (rc) goto done;
  rc = og_toff(g, bn);
done:
  og_field_free(rbr); og_field_free(rbp); og_field_free(bz);
  og_field_free(psi); og_field_free(chi); og_field_free(w);
  return rc;
}

/* CALC_BOUSSI_DIAG + CALC_BOUSSI_ETD_OP: mod_boussinesq.f90:42-109,
 * mod_march.f90:906-945.  Returns the real parts (col-major 4x4). */
void og_etd_op(const og_grid *g, double dt, double *e16, double *nl16) {
  double nf = g->p.bv0, om = g->p.omega;
  cplx jd[4] = {-I * nf, I * nf, -2.0 * I * om, 2.0 * I * om};
  cplx s[16] = {0}, si[16] = {0};
#define M4(a, i, j) a[(i) + 4 * (j)]
  M4(s, 0, 2) = I; M4(s, 0, 3) = -I; M4(s, 1, 2) = 1.0; M4(s, 1, 3) = 1.0;
  M4(s, 2, 0) = -I / nf; M4(s, 2, 1) = I / nf; M4(s, 3, 0) = 1.0; M4(s, 3, 1) = 1.0;
  M4(si, 0, 2) = I * nf / 2.0; M4(si, 0, 3) = 0.5; M4(si, 1, 2) = -I * nf / 2.0; M4(si, 1, 3) = 0.5;
  M4(si, 2, 0) = -I / 2.0; M4(si, 2, 1) = 0.5; M4(si, 3, 0) = I / 2.0; M4(si, 3, 1) = 0.5;
  cplx ex[4], ph[4];
  for (int i = 0; i < 4; i++) {
    ex[i] = cexp(jd[i] * dt);
    ph[i] = (cabs(jd[i]) > 1.0e-14) ? (ex[i] - 1.0) / jd[i] : dt;
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      cplx ae = 0, an = 0;
      for (int l = 0; l < 4; l++) {
        ae += M4(s, i, l) * ex[l] * M4(si, l, j);
        an += M4(s, i, l) * ph[l] * M4(si, l, j);
      }
      e16[i + 4 * j] = creal(ae);
      nl16[i + 4 * j] = creal(an);
    }
#undef M4
}

/* PRODCT_MK_HALFK / L2NORM_MK_HALFK: mod_scalar3.f90:470-554,750-841 */
static void prodct_halfk(const og_grid *g, const og_field *a, const og_field *b, int l2, double *out) {
  int nt = g->ntchop, nxc = g->nxchop, nxd = g->nxchopdim;
  double *all = (double *)calloc((size_t)nt * nxd, sizeof(double));
  const double pi = acos(-1.0);
  for (int kk = 0; kk < nxd; kk++)
    for (int mm = 0; mm < g->ntchopdim && mm < nt; mm++) {
      double s = 0;
      for (int r = 0; r < g->nr; r++) {
        cplx av = a->e[QIDX(g, r, mm, kk)];
        double wk;
        if (l2) {
          double ab = cabs(av);
          wk = pow(ab, 2.0) / pow(1 - g->x[r], 2.0);
        } else {
          wk = creal(av * conj(b->e[QIDX(g, r, mm, kk)]));
        }
        s += wk * g->w[r];
      }
      all[mm + (size_t)nt * kk] = 4.0 * pi * g->zlen0 * g->ell2 * s;
    }
  for (int kk = 0; kk < nxd; kk++) all[(size_t)nt * kk] = 0.5 * all[(size_t)nt * kk];
  for (int i = 0; i < nt * nxc; i++) out[i] = 0;
  for (int mm = 0; mm < nt; mm++) out[mm] = all[mm];
  for (int kk = 2; kk <= nxc; kk++) {
    int kc = nxd + 2 - kk;
    for (int mm = 0; mm < nt; mm++) out[mm + (size_t)nt * (kk - 1)] = all[mm + (size_t)nt * (kc - 1)] + all[mm + (size_t)nt * (kk - 1)];
  }
  free(all);
}

/* CALC_BOUSSI_ENERGY: mod_boussinesq.f90:138-232 (file output omitted) */
int og_energy(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b, double *ke, double *pe, double *dke) {
  if (b->ln != 0.0) return E_LOGTERM;
  int nt = g->ntchop, nxc = g->nxchop;
  og_field *w1 = og_field_new(g, OG_FFF), *w2 = og_field_new(g, OG_FFF), *w3 = og_field_new(g, OG_FFF);
  double *tmp = (double *)calloc((size_t)nt * nxc, sizeof(double));
  og_chopset(g, 3);
  og_copy(g, w1, psi);
  w1->ln = 2.0 * psi->ln;
  og_delsqh(g, w1, w2);
  w1->ln = 0.0; w2->ln = 0.0;
  og_rtran(g, w1, 1); og_rtran(g, w2, 1);
  prodct_halfk(g, w1, w2, 0, tmp);
  for (int i = 0; i < nt * nxc; i++) ke[i] = -tmp[i];
  realloc_space(g, w1, OG_FFF); realloc_space(g, w2, OG_FFF);
  int rc = og_del2(g, chi, w1);
  og_delsqh(g, chi, w2);
  w1->ln = 0.0; w2->ln = 0.0;
  og_rtran(g, w1, 1); og_rtran(g, w2, 1);
  prodct_halfk(g, w1, w2, 0, tmp);
  for (int i = 0; i < nt * nxc; i++) ke[i] = (mutation == 4 ? ke[i] - tmp[i] : ke[i] + tmp[i]) / 2.0;
  og_copy(g, w3, b);
  w3->ln = 0.0;
  og_rtran(g, w3, 1);
  prodct_halfk(g, w3, w2, 0, dke);
  if (mutation == 5)
    for (int i = 0; i < nt * nxc; i++) dke[i] = -dke[i];
  prodct_halfk(g, w3, w3, 1, tmp);
  for (int i = 0; i < nt * nxc; i++) pe[i] = tmp[i] / 2.0 / pow(g->p.bv0, 2.0);
  og_chopset(g, -3);
  og_field_free(w1); og_field_free(w2); og_field_free(w3);
  free(tmp);
  return rc;
}

/* ===================================================================== */
/* Exponential hyperviscosity filter: HYP3SET/HYP3MASK/HYP3INIT           */
/* mod_march.f90:3320-3480                                                */
/* ===================================================================== */
static void hyp3mask(const double *kv, int n, double ratio, double steep, double *mask) {
  double kmax = -1e300;
  for (int i = 0; i < n; i++) kmax = kv[i] > kmax ? kv[i] : kmax;
  if (kmax < 1.0e-12) kmax = 1.0;
  double ta = tanh(steep);
  for (int i = 0; i < n; i++) {
    double kn = kv[i] / kmax;
    if (kn < ratio) mask[i] = 0.0;
    else {
      double xr = (kn - ratio) / (1.0 - ratio);
      double yr = steep * (2.0 * xr - 1.0);
      mask[i] = (tanh(yr) + ta) / (2.0 * ta);
    }
  }
}

static void hyp3set(og_grid *g, double nu_r, double nu_t, double nu_x, double **hr, double **ht, double **hx) {
  int nrc = g->p.nrchop, ntd = g->ntchopdim, nxd = g->nxchopdim, ph = g->p.visc_p / 2;
  free(*hr); free(*ht); free(*hx);
  *hr = (double *)calloc((size_t)nrc * ntd, sizeof(double));
  *ht = (double *)calloc((size_t)ntd, sizeof(double));
  *hx = (double *)calloc((size_t)nxd, sizeof(double));
  double *mr = (double *)calloc((size_t)nrc * ntd, sizeof(double)), *mt = (double *)calloc((size_t)ntd, sizeof(double));
  double *mx = (double *)calloc((size_t)nxd, sizeof(double)), *kv = (double *)calloc((size_t)(nrc > nxd ? nrc : nxd) + ntd, sizeof(double));
  for (int mm = 0; mm < ntd; mm++) {
    int nn = imax(imin(nrc, nrc - g->m[mm]), 0);
    for (int i = 0; i < nrc; i++) mr[i + (size_t)nrc * mm] = 1.0;
    for (int i = 0; i < nn; i++) kv[i] = (double)i;
    if (nn > 0) hyp3mask(kv, nn, 0.5, 2.0, mr + (size_t)nrc * mm);
  }
  for (int mm = 0; mm < ntd; mm++) kv[mm] = (double)g->m[mm];
  hyp3mask(kv, ntd, 2.0 / 3.0, 2.0, mt);
  for (int kk = 0; kk < nxd; kk++) kv[kk] = fabs(AK(g, 0, kk));
  hyp3mask(kv, nxd, 2.0 / 3.0, 2.0, mx);
  double k2r = 1.5 * 1.5 * (double)((nrc - 1) * nrc) / g->ell2;
  double k2t = 0, k2x = 0;
  for (int mm = 0; mm < ntd; mm++) k2t = fmax(k2t, (double)(g->m[mm] * g->m[mm]));
  for (int kk = 0; kk < nxd; kk++) k2x = fmax(k2x, AK(g, 0, kk) * AK(g, 0, kk));
  for (int mm = 0; mm < ntd; mm++) {
    (*ht)[mm] = exp(-nu_t * mt[mm] * pow((double)(g->m[mm] * g->m[mm]) / k2t, ph));
    for (int nn = 0; nn < nrc; nn++)
      (*hr)[nn + (size_t)nrc * mm] = exp(-nu_r * mr[nn + (size_t)nrc * mm] * pow(1.5 * 1.5 * (double)(nn * (nn + 1)) / g->ell2 / k2r, ph));
  }
  for (int kk = 0; kk < nxd; kk++) (*hx)[kk] = exp(-nu_x * mx[kk] * pow(AK(g, 0, kk) * AK(g, 0, kk) / k2x, ph));
  free(mr); free(mt); free(mx); free(kv);
}

/* HYP3INIT (mod_march.f90:3463-3480): every direction at NUP, KAPPAP */
static void hyp3_init(og_grid *g) {
  hyp3set(g, g->nup_r, g->nup_t, g->nup_x, &g->hv_r, &g->hv_t, &g->hv_x);
  hyp3set(g, g->kap_r, g->kap_t, g->kap_x, &g->hd_r, &g->hd_t, &g->hd_x);
}

/* multiply n < NRCHOPS(m) by (H_r H_t H_x)^tf: HYPERV3 / CALC_BOUSSI_ETD_HYPERV */
static void hyp3_apply(const og_grid *g, og_field *a, int diff, double tf) {
  const double *hr = diff ? g->hd_r : g->hv_r, *ht = diff ? g->hd_t : g->hv_t, *hx = diff ? g->hd_x : g->hv_x;
  int nrc = g->p.nrchop;
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++)
      for (int n = 0; n < g->nrchops[mm]; n++) {
        double h = hr[n + (size_t)nrc * mm] * ht[mm] * hx[kk];
        a->e[FIDX(g, n, mm, kk)] = a->e[FIDX(g, n, mm, kk)] * pow(h, tf);
      }
  og_chopdo(g, a);
}

static int filter_active(const og_grid *g) { return g->p.visc_sw == 3 || g->p.hyper_cn; }

/* ===================================================================== */
/* Vorticity removal, axial sponge and adaptive hyperviscosity            */
/* ===================================================================== */

/* TOPHAT: mod_march.f90:3704-3712 */
static double tophat(double x, double hw, double st) { return 0.5 * (tanh((x + hw) / st) - tanh((x - hw) / st)); }

/* SPNGINIT: mod_march.f90:2292-2323 (N_TAU = 10, RATIO_L = 0.86,
 * RATIO_S = 0.04; the reference compiles default reals as REAL(8)).
 * rmv_sw != 2 turns the sponge off. */
void og_set_sponge(og_grid *g, int rmv_sw) {
  g->rmv_sw = rmv_sw;
  free(g->tophat_x);
  g->tophat_x = NULL;
  if (rmv_sw != 2) return;
  const double ratio_l = 0.86, ratio_s = 0.04;
  const double hlfz = g->zlen / 2.0;
  g->sp_tau = 10 * g->p.dt;
  const double hw = g->zlen * ratio_l / 2.0, st = g->zlen * ratio_s;
  g->tophat_x = (double *)calloc((size_t)g->ndimx, sizeof(double));
  for (int k = 0; k < g->nx; k++) g->tophat_x[k] = tophat(g->z[k] - hlfz, hw, st);
}

/* the sponge's damping of every z plane over one application of length dt */
static void sponge_damp(const og_grid *g, double dt, double *damp) {
  for (int k = 0; k < g->ndimx; k++) damp[k] = exp(-(1.0 - g->tophat_x[k]) / g->sp_tau * dt);
}

/* PFP field (r, m < NTCHOPDIM in the NDIMTH slots, z): z plane k times
 * damp[k], then m = 0 made real */
static void sponge_planes(const og_grid *g, og_field *f, const double *damp) {
  for (int z = 0; z < g->ndimx; z++)
    for (int mm = 0; mm < g->ntchopdim; mm++)
      for (int r = 0; r < g->ndimr; r++) f->e[PIDX(g, r, mm, z)] = f->e[PIDX(g, r, mm, z)] * damp[z];
  for (int z = 0; z < g->ndimx; z++)
    for (int r = 0; r < g->ndimr; r++) f->e[PIDX(g, r, 0, z)] = creal(f->e[PIDX(g, r, 0, z)]);
}

/* CALC_BOUSSI_ETD_SPONGE: mod_march.f90:1436-1469 */
static int etd_sponge(og_grid *g, og_field *uz, og_field *b, double dt) {
  if (g->rmv_sw != 2) return E_OK;
  int rc = E_OK;
  double *damp = (double *)calloc((size_t)g->ndimx, sizeof(double));
  sponge_damp(g, dt, damp);
  og_field *fs[2] = {uz, b};
  for (int i = 0; i < 2 && !rc; i++) {
    rc = og_rtran(g, fs[i], 1);
    if (!rc) rc = og_verfft(g, fs[i], 1);
  }
  for (int i = 0; i < 2 && !rc; i++) sponge_planes(g, fs[i], damp);
  for (int i = 0; i < 2 && !rc; i++) rc = og_toff(g, fs[i]);
  free(damp);
  return rc;
}

/* SPONGE: mod_march.f90:2380-2420 (b may be NULL) */
int og_sponge(og_grid *g, og_field *psi, og_field *chi, og_field *b, double dt) {
  if (g->rmv_sw != 2) return E_OK;
  if (psi->space != OG_FFF) return E_TRANSFORM;
  int rc = E_OK;
  const double psi_ln = psi->ln, chi_ln = chi->ln;
  psi->ln = 0.0;
  chi->ln = 0.0;
  double *damp = (double *)calloc((size_t)g->ndimx, sizeof(double));
  sponge_damp(g, dt, damp);
  og_field *fs[3] = {psi, chi, b};
  const int nf = b ? 3 : 2;
  for (int i = 0; i < nf && !rc; i++) {
    rc = og_rtran(g, fs[i], 1);
    if (!rc) rc = og_verfft(g, fs[i], 1);
  }
  for (int i = 0; i < nf && !rc; i++) sponge_planes(g, fs[i], damp);
  for (int i = 0; i < nf && !rc; i++) rc = og_toff(g, fs[i]);
  psi->ln = psi_ln;
  chi->ln = chi_ln;
  free(damp);
  return rc;
}

/* SMOOTH: mod_march.f90:3562-3600 (ni >= 3); t: ni scratch values */
static void smooth(cplx *a, int ni, cplx *t) {
  static const double fae[3] = {0.2, 0.3, 0.5};
  for (int i = 0; i < ni; i++) t[i] = 0;
  t[0] = a[0];
  for (int j = 0; j < 3; j++) t[ni - 3 + j] = t[ni - 3 + j] + a[ni - 1] * fae[j];
  for (int i = 2; i <= ni - 1; i++) { /* the reference's 1-based index */
    const double f = (1 + ((ni - i) / (ni - 1.0))) * 0.5;
    t[i - 2] = t[i - 2] + a[i - 1] * (1 - f) / 2;
    t[i - 1] = t[i - 1] + a[i - 1] * f;
    t[i] = t[i] + a[i - 1] * (1 - f) / 2;
  }
  memcpy(a, t, sizeof(cplx) * (size_t)ni);
}

/* SMOOTH applied `passes` times to ni interleaved complex values (tests) */
int og_smooth(double *v, int ni, int passes) {
  if (ni < 3) return E_CALC;
  cplx *a = (cplx *)calloc((size_t)ni, sizeof(cplx)), *t = (cplx *)calloc((size_t)ni, sizeof(cplx));
  for (int i = 0; i < ni; i++) a[i] = v[2 * i] + I * v[2 * i + 1];
  for (int p = 0; p < passes; p++) smooth(a, ni, t);
  for (int i = 0; i < ni; i++) {
    v[2 * i] = creal(a[i]);
    v[2 * i + 1] = cimag(a[i]);
  }
  free(a);
  free(t);
  return E_OK;
}

/* RINGCUT: mod_march.f90:2403-2445, on a PFF field; ns0, ns are the
 * reference's 1-based radial indices.  Rows ns0..NR are zeroed except at
 * (m, k) = (0, 0); then rows ns..NR of every (m, k) column are weighted by
 * w (divxm2: w / (1-x)^2), smoothed 5 times and unweighted. */
int og_ringcut(og_grid *g, og_field *wk, int ns0, int ns, int divxm2) {
  const int nr = g->nr, ni = nr - ns + 1;
  if (wk->space != OG_PFF || ns < 1 || ns0 < 1 || ni < 3) return E_TRANSFORM;
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      if (mm == 0 && (kk == 0 || g->nx <= 1)) continue;
      for (int r = ns0 - 1; r < nr; r++) wk->e[QIDX(g, r, mm, kk)] = 0;
    }
  double *wf = (double *)calloc((size_t)ni, sizeof(double));
  for (int i = 0; i < ni; i++) {
    const int r = ns - 1 + i;
    wf[i] = divxm2 ? g->w[r] / ((1.0 - g->x[r]) * (1.0 - g->x[r])) : g->w[r];
  }
#pragma omp parallel for collapse(2)
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      cplx t[ni], *col = wk->e + QIDX(g, ns - 1, mm, kk);
      for (int i = 0; i < ni; i++) col[i] = col[i] * wf[i];
      for (int p = 0; p < 5; p++) smooth(col, ni, t);
      for (int i = 0; i < ni; i++) col[i] = col[i] / wf[i];
    }
  free(wf);
  return E_OK;
}

/* REMOVE: mod_march.f90:2503-2551 (b may be NULL) */
int og_remove(og_grid *g, og_field *psi, og_field *chi, og_field *b) {
  const int ns = g->nr * 2 / 3, ns0 = imin(ns + 4, g->nr);
  og_field *wk = og_field_new(g, OG_FFF);
  int rc = og_delsqh(g, psi, wk);
  if (!rc) rc = og_rtran(g, wk, 1);
  if (!rc) rc = og_ringcut(g, wk, ns0, ns, 0);
  if (!rc) rc = og_rtran(g, wk, -1);
  if (!rc) rc = og_idelsqh(g, wk, psi);
  if (!rc) {
    og_chopset(g, 2);
    realloc_space(g, wk, OG_FFF);
    rc = og_del2(g, chi, wk);
    if (!rc) rc = og_rtran(g, wk, 1);
    if (!rc) rc = og_ringcut(g, wk, ns0, ns, 1);
    og_chopset(g, -2);
    if (!rc) rc = og_rtran(g, wk, -1);
    const double ln = chi->ln;
    if (!rc) rc = og_idel2(g, wk, chi, 1, ln);
  }
  og_field_free(wk);
  if (!rc && b) {
    rc = og_rtran(g, b, 1);
    if (!rc) rc = og_ringcut(g, b, ns0, ns, 1);
    if (!rc) rc = og_rtran(g, b, -1);
  }
  return rc;
}

/* ENEMON: mod_march.f90:2581-2617 at the caller's truncation; em[NTCHOP],
 * ek[NXCHOP] */
int og_enemon(og_grid *g, const og_field *psi, const og_field *chi, double *em, double *ek) {
  const int nt = g->ntchop, nxc = g->nxchop;
  og_field *w1 = og_field_new(g, OG_FFF), *w2 = og_field_new(g, OG_FFF);
  double *tmp = (double *)calloc((size_t)nt * nxc, sizeof(double));
  og_copy(g, w1, psi);
  w1->ln = 2 * psi->ln;
  int rc = og_delsqh(g, w1, w2);
  if (!rc) rc = og_rtran(g, w2, 1);
  og_copy(g, w1, psi);
  w1->ln = 0.0;
  if (!rc) rc = og_rtran(g, w1, 1);
  if (!rc) {
    prodct_halfk(g, w2, w1, 0, tmp);
    for (int mm = 0; mm < nt; mm++) {
      double a = 0.0;
      for (int kk = 0; kk < nxc; kk++) a = a + tmp[mm + (size_t)nt * kk];
      em[mm] = a;
    }
    for (int kk = 0; kk < nxc; kk++) {
      double a = 0.0;
      for (int mm = 0; mm < nt; mm++) a = a + tmp[mm + (size_t)nt * kk];
      ek[kk] = a;
    }
    realloc_space(g, w1, OG_FFF);
    realloc_space(g, w2, OG_FFF);
    rc = og_del2(g, chi, w1);
  }
  if (!rc) rc = og_rtran(g, w1, 1);
  if (!rc) rc = og_delsqh(g, chi, w2);
  if (!rc) rc = og_rtran(g, w2, 1);
  if (!rc) {
    prodct_halfk(g, w2, w1, 0, tmp);
    for (int mm = 0; mm < nt; mm++) {
      double a = 0.0;
      for (int kk = 0; kk < nxc; kk++) a = a + tmp[mm + (size_t)nt * kk];
      em[mm] = -em[mm] + a;
    }
    for (int kk = 0; kk < nxc; kk++) {
      double a = 0.0;
      for (int mm = 0; mm < nt; mm++) a = a + tmp[mm + (size_t)nt * kk];
      ek[kk] = -ek[kk] + a;
    }
  }
  og_field_free(w1);
  og_field_free(w2);
  free(tmp);
  return rc;
}

/* SPECURV: mod_march.f90:2642-2689; espec[0..n-1] */
double og_specurv(const double *espec, int n, double ratio) {
  const double floor_ = 1.0e-22, target = -0.05;
  int ie = n, is = ie - imax(1, (int)(ie * ratio));
  while (ie > is && espec[ie - 1] < floor_) ie--;
  const double rtrunc = 1.0 - (double)(ie - is) / (double)(n - is);
  const int nn = ie - is + 1;
  if (nn < 10 || rtrunc > 0.2) return target;
  double curv = 0.0;
  const double dx = 2.0 / (nn - 1);
  for (int i = 1; i <= nn; i++) {
    const double x = -1.0 + 2.0 * (i - 1) / (nn - 1);
    const double fac = (i == 1 || i == nn) ? 0.5 : 1.0;
    curv = curv + fac * log10(espec[is + i - 2]) * dx * sqrt(5.0 / 8.0) * (3.0 * (x * x) - 1.0);
  }
  return fmin(10.0, fmax(-10.0, curv));
}

/* SPESLOP: mod_march.f90:2729-2800; espec[0..n-1] */
double og_speslop(const double *espec, int n, double ratio) {
  const double floor_ = 1.0e-24, drop = 1.0e10;
  int ie = n, is = ie - imax(1, (int)(ie * ratio));
  const int k_start = is;
  double peak = espec[is - 1];
  for (int i = is; i <= ie; i++) peak = fmax(peak, espec[i - 1]);
  const double rel = peak / drop;
  while (ie > is && (espec[ie - 1] < floor_ || espec[ie - 1] < rel)) ie--;
  const int nn = ie - is + 1;
  if (nn < 3) return -3.0;
  const double xi_min = log10((double)k_start), xi_max = log10((double)(k_start + nn - 1));
  double slope = 0.0;
  const double dx = 2.0 / (nn - 1);
  for (int i = 1; i <= nn; i++) {
    const double xi = log10((double)(k_start + i - 1));
    const double x = -1.0 + 2.0 * (xi - xi_min) / (xi_max - xi_min);
    const double fac = (i == 1 || i == nn) ? 0.5 : 1.0;
    slope = slope + fac * log10(espec[is + i - 2]) * dx * sqrt(3.0 / 2.0) * x;
  }
  slope = slope * 2.0 / (xi_max - xi_min);
  return fmin(5.0, fmax(-15.0, slope));
}

/* PDAJUST: mod_march.f90:2841-2867; params = [KP, KD, OFFSET, SCALE];
 * has_thr: the over-damping threshold is given */
void og_pdajust(double *coeff, double cur, double prev, double base, const double *params, int has_thr, double thr) {
  double control = params[0] * (cur - prev) + params[1] * cur + params[2];
  if (has_thr && cur < thr && prev < thr) {
    const double severity = fmin(5.0, fabs(cur - thr));
    control = -(0.1 + 0.05 * severity);
  }
  control = fmin(1.0, fmax(-0.4, params[3] * control));
  *coeff = (1.0 + control) * *coeff;
  if (*coeff < base) *coeff = base;
}

/* the dual controllers of one coefficient set (HYP3ADJ's velocity or
 * density part): c[0..2] = (R, T, X) coefficients, hist[0..3] = (CURVMO,
 * CURVKO, SLOPMO, SLOPKO), base the baseline */
static void hyp3_controllers(double *c, double *hist, double base, double smc, double skc, double sms, double sks) {
  static const double cp[4] = {2.0, 1.0, 0.05, 0.8}, sp[4] = {1.5, 0.5, 1.5, 1.0};
  double tc = c[1], xc = c[2], ts = c[1], xs = c[2];
  og_pdajust(&tc, smc, hist[0], base, cp, 1, -1.0);
  og_pdajust(&xc, skc, hist[1], base, cp, 1, -1.0);
  og_pdajust(&ts, sms, hist[2], base, sp, 0, 0.0);
  og_pdajust(&xs, sks, hist[3], base, sp, 0, 0.0);
  c[1] = fmax(tc, ts);
  c[2] = fmax(xc, xs);
  double rco, rcn, rso, rsn;
  if (smc > skc) {
    rco = (3.0 * hist[0] + hist[1]) / 4.0;
    rcn = (3.0 * smc + skc) / 4.0;
  } else {
    rco = (hist[0] + 3.0 * hist[1]) / 4.0;
    rcn = (smc + 3.0 * skc) / 4.0;
  }
  double rc = c[0];
  og_pdajust(&rc, rcn, rco, base, cp, 1, -1.0);
  if (sms > sks) {
    rso = (3.0 * hist[2] + hist[3]) / 4.0;
    rsn = (3.0 * sms + sks) / 4.0;
  } else {
    rso = (hist[2] + 3.0 * hist[3]) / 4.0;
    rsn = (sms + 3.0 * sks) / 4.0;
  }
  double rs = c[0];
  og_pdajust(&rs, rsn, rso, base, sp, 0, 0.0);
  c[0] = fmin(fmax(rc, rs), 2.0);
  hist[0] = smc;
  hist[1] = skc;
  hist[2] = sms;
  hist[3] = sks;
}

/* HYP3ADJ (mod_march.f90:3000-3182) followed by the two HYP3SETs that
 * TIME_STEPPING makes after it (:282-287); b may be NULL */
int og_hyp3adj(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b) {
  const int nt = g->ntchop, nxc = g->nxchop;
  double *em = (double *)calloc((size_t)nt, sizeof(double)), *ek = (double *)calloc((size_t)nxc, sizeof(double));
  og_chopset(g, 2);
  int rc = og_enemon(g, psi, chi, em, ek);
  og_chopset(g, -2);
  if (!rc) {
    const double smc = og_specurv(em, nt - 2, 1.0 / 4.0), skc = og_specurv(ek, nxc, 1.0 / 4.0);
    const double sms = og_speslop(em, nt - 2, 1.0 / 3.0), sks = og_speslop(ek, nxc, 1.0 / 3.0);
    if (g->nupb < 0.0) {
      g->curvmo = smc;
      g->curvko = skc;
      g->slopmo = sms;
      g->slopko = sks;
      g->nupb = 0.8;
    }
    double c[3] = {g->nup_r, g->nup_t, g->nup_x}, h[4] = {g->curvmo, g->curvko, g->slopmo, g->slopko};
    hyp3_controllers(c, h, g->nupb, smc, skc, sms, sks);
    g->nup_r = c[0]; g->nup_t = c[1]; g->nup_x = c[2];
    g->curvmo = h[0]; g->curvko = h[1]; g->slopmo = h[2]; g->slopko = h[3];
  }
  if (!rc && b) {
    /* CALC_BOUSSI_PE (mod_boussinesq.f90:112-135) at the current truncation */
    og_field *w = og_field_new(g, OG_FFF);
    double *pe = (double *)calloc((size_t)nt * nxc, sizeof(double));
    double *pm = (double *)calloc((size_t)nt, sizeof(double)), *pk = (double *)calloc((size_t)nxc, sizeof(double));
    og_copy(g, w, b);
    w->ln = 0.0;
    rc = og_rtran(g, w, 1);
    if (!rc) {
      prodct_halfk(g, w, w, 1, pe);
      for (int i = 0; i < nt * nxc; i++) pe[i] = pe[i] / 2.0 / (g->p.bv0 * g->p.bv0);
      for (int mm = 0; mm < nt; mm++) {
        double a = 0.0;
        for (int kk = 0; kk < nxc; kk++) a = a + pe[mm + (size_t)nt * kk];
        pm[mm] = a;
      }
      for (int kk = 0; kk < nxc; kk++) {
        double a = 0.0;
        for (int mm = 0; mm < nt; mm++) a = a + pe[mm + (size_t)nt * kk];
        pk[kk] = a;
      }
      const double smc = og_specurv(pm, nt - 2, 1.0 / 4.0), skc = og_specurv(pk, nxc, 1.0 / 4.0);
      const double sms = og_speslop(pm, nt - 2, 1.0 / 3.0), sks = og_speslop(pk, nxc, 1.0 / 3.0);
      if (g->kappapb < 0.0) {
        g->pcurvmo = smc;
        g->pcurvko = skc;
        g->pslopmo = sms;
        g->pslopko = sks;
        g->kappapb = g->nupb;
      }
      double c[3] = {g->kap_r, g->kap_t, g->kap_x}, h[4] = {g->pcurvmo, g->pcurvko, g->pslopmo, g->pslopko};
      hyp3_controllers(c, h, g->kappapb, smc, skc, sms, sks);
      g->kap_r = c[0]; g->kap_t = c[1]; g->kap_x = c[2];
      g->pcurvmo = h[0]; g->pcurvko = h[1]; g->pslopmo = h[2]; g->pslopko = h[3];
    }
    og_field_free(w);
    free(pe); free(pm); free(pk);
  }
  if (!rc) hyp3_init(g);
  free(em);
  free(ek);
  return rc;
}

/* the six directional coefficients (NUP_R, NUP_T, NUP_X, KAPPAP_R,
 * KAPPAP_T, KAPPAP_X); set also rebuilds the filters (HYP3SET) */
void og_hyperv_coeffs(const og_grid *g, double *out) {
  out[0] = g->nup_r; out[1] = g->nup_t; out[2] = g->nup_x;
  out[3] = g->kap_r; out[4] = g->kap_t; out[5] = g->kap_x;
}
void og_set_hyperv_coeffs(og_grid *g, const double *in) {
  g->nup_r = in[0]; g->nup_t = in[1]; g->nup_x = in[2];
  g->kap_r = in[3]; g->kap_t = in[4]; g->kap_x = in[5];
  hyp3_init(g);
}

/* ===================================================================== */
/* Time march: mod_march.f90                                              */
/* ===================================================================== */
og_solver *og_solver_new(og_grid *g) {
  og_solver *s = (og_solver *)calloc(1, sizeof(og_solver));
  s->g = g;
  s->n_psi = og_field_new(g, OG_FFF); s->n_chi = og_field_new(g, OG_FFF); s->n_b = og_field_new(g, OG_FFF);
  s->r_n_r = og_field_new(g, OG_FFF); s->r_n_p = og_field_new(g, OG_FFF); s->n_z = og_field_new(g, OG_FFF);
  s->r_pi_r = og_field_new(g, OG_FFF); s->r_pi_p = og_field_new(g, OG_FFF); s->pi_z = og_field_new(g, OG_FFF);
  return s;
}

void og_solver_free(og_solver *s) {
  if (!s) return;
  og_field_free(s->n_psi); og_field_free(s->n_chi); og_field_free(s->n_b);
  og_field_free(s->r_n_r); og_field_free(s->r_n_p); og_field_free(s->n_z);
  og_field_free(s->r_pi_r); og_field_free(s->r_pi_p); og_field_free(s->pi_z);
  free(s);
}

double og_time(const og_solver *s) { return s->t; }
int og_nstep(const og_solver *s) { return s->n; }

/* VISC1: mod_march.f90:1529-1576 (backward Euler) */
static int visc1(og_grid *g, og_field *a, double nu, double dt, og_field *wk) {
  if (nu == 0.0) return E_SOLVER;
  double alp = 1 / (nu * dt);
  realloc_space(g, wk, OG_FFF);
  wk->ln = -alp * a->ln;
  long n = og_field_size(g, OG_FFF);
  for (long i = 0; i < n; i++) wk->e[i] = -alp * a->e[i];
  return og_ihelm(g, wk, a, alp);
}

/* VISC2: mod_march.f90:1579-1636 (Crank-Nicolson) */
static int visc2(og_grid *g, og_field *a, const og_field *ao, double nu, double dt, og_field *wk) {
  if (nu == 0.0) return E_SOLVER;
  double alp = 2 / nu / dt;
  og_chopset(g, 2);
  int rc = og_del2(g, ao, wk);
  og_chopset(g, -2);
  if (rc) return rc;
  og_chopdo(g, wk);
  long n = og_field_size(g, OG_FFF);
  for (long i = 0; i < n; i++) wk->e[i] = -wk->e[i] - alp * a->e[i];
  wk->ln = -alp * a->ln;
  return og_ihelm(g, wk, a, alp);
}

/* CALC_BOUSSI_VISC_BE: mod_march.f90:1200-1230 */
static int visc_be(og_grid *g, og_field *psi, og_field *chi, og_field *b, double dt) {
  if (g->p.visc_sw != 1) return E_OK;
  og_field *wk = og_field_new(g, OG_FFF);
  int rc = visc1(g, psi, g->p.nu, dt, wk);
  if (!rc) rc = visc1(g, chi, g->p.nu, dt, wk);
  if (!rc && g->p.kappa != 0.0) rc = visc1(g, b, g->p.kappa, dt, wk);
  og_field_free(wk);
  return rc;
}

/* CALC_BOUSSI_VISC_CN: mod_march.f90:1233-1265 */
static int visc_cn(og_grid *g, og_field *psi, og_field *chi, og_field *b, const og_field *psio, const og_field *chio,
                   const og_field *bo) {
  if (g->p.visc_sw != 1) return E_OK;
  og_field *wk = og_field_new(g, OG_FFF);
  int rc = visc2(g, psi, psio, g->p.nu, g->p.dt, wk);
  if (!rc) rc = visc2(g, chi, chio, g->p.nu, g->p.dt, wk);
  if (!rc && g->p.kappa != 0.0) rc = visc2(g, b, bo, g->p.kappa, g->p.dt, wk);
  og_field_free(wk);
  return rc;
}

/* HYPERV: mod_march.f90:1639-1686, implicit Euler step of
 * dA/dt = nu del^2 A -(-1)^(P/2) nup del^P A through IHELMP. */
static int hyperv2(og_grid *g, og_field *a, double nu, double nup, int p, double dt) {
  if (nup == 0.0) return E_SOLVER;
  if (a->space != OG_FFF) return E_SOLVER;
  const int ph = p / 2;
  const double sg = (ph & 1) ? -1.0 : 1.0;
  const double alp = 1 / dt / nup * sg, bet = -sg * nu / nup;
  og_field *wk = og_field_new(g, OG_FFF);
  for (size_t i = 0; i < (size_t)g->nrchopdim * g->ntchopdim * g->nxchopdim; i++) wk->e[i] = alp * a->e[i];
  wk->ln = alp * a->ln;
  int rc = og_ihelmp(g, p, wk, a, -alp, bet);
  og_field_free(wk);
  return rc;
}

/* CALC_BOUSSI_HYPERV: mod_march.f90:1305-1349 (SW=2: HYPERV of psi, chi and,
 * with a hyperdiffusivity, b; SW=3: the exponential filter) */
static int hyperv_state(og_grid *g, og_field *psi, og_field *chi, og_field *b, double dt) {
  if (g->p.visc_sw == 2) {
    int rc = hyperv2(g, psi, g->p.nu, g->p.nup, g->p.visc_p, dt);
    if (!rc) rc = hyperv2(g, chi, g->p.nu, g->p.nup, g->p.visc_p, dt);
    if (!rc && g->p.kappap != 0.0) rc = hyperv2(g, b, g->p.kappa, g->p.kappap, g->p.visc_p, dt);
    return rc;
  }
  if (g->p.visc_sw != 3) return E_OK;
  double tf = dt / g->p.dt;
  hyp3_apply(g, psi, 0, tf);
  hyp3_apply(g, chi, 0, tf);
  hyp3_apply(g, b, 1, tf);
  return E_OK;
}

/* CALC_BOUSSI_ETD_HYPERV: mod_march.f90:1352-1416 (SW=3; with hyper_cn the
 * same filter also runs beside CN viscosity) */
static void etd_hyperv(og_grid *g, og_field *rur, og_field *rup, og_field *uz, og_field *b, double dt) {
  if (!filter_active(g)) return;
  double tf = dt / g->p.dt;
  hyp3_apply(g, rur, 0, tf);
  hyp3_apply(g, rup, 0, tf);
  hyp3_apply(g, uz, 0, tf);
  hyp3_apply(g, b, 1, tf);
}

static void axpy3(og_grid *g, og_field *y, const og_field *x, double a) { scale_add(g, y, x, a, OG_FFF); y->ln = y->ln + a * x->ln; }

/* RICH_BOUSSI_FE_BE: mod_march.f90:535-606 */
static int rich_fe_be(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  double dt = g->p.dt, hdt = dt * 0.5;
  int rc;
  s->t += dt;
  s->n += 1;
  og_field *psi2 = og_field_new(g, OG_FFF), *chi2 = og_field_new(g, OG_FFF), *b2 = og_field_new(g, OG_FFF);
  og_field *pn = og_field_new(g, OG_FFF), *cn = og_field_new(g, OG_FFF), *bnn = og_field_new(g, OG_FFF);
  long n = og_field_size(g, OG_FFF);
  if ((rc = og_force(g, psi, chi, b, s->n_psi, s->n_chi, s->n_b))) goto done;
  for (long i = 0; i < n; i++) {
    psi2->e[i] = psi->e[i] + (hdt * s->n_psi->e[i]);
    chi2->e[i] = chi->e[i] + (hdt * s->n_chi->e[i]);
    b2->e[i] = b->e[i] + (hdt * s->n_b->e[i]);
  }
  psi2->ln = psi->ln + (hdt * s->n_psi->ln);
  chi2->ln = chi->ln + (hdt * s->n_chi->ln);
  if ((rc = visc_be(g, psi2, chi2, b2, hdt))) goto done;
  hyperv_state(g, psi2, chi2, b2, hdt);
  if ((rc = og_force(g, psi2, chi2, b2, pn, cn, bnn))) goto done;
  for (long i = 0; i < n; i++) {
    psi2->e[i] = psi2->e[i] + (hdt * pn->e[i]);
    chi2->e[i] = chi2->e[i] + (hdt * cn->e[i]);
    b2->e[i] = b2->e[i] + (hdt * bnn->e[i]);
  }
  psi2->ln = psi2->ln + (hdt * pn->ln);
  chi2->ln = chi2->ln + (hdt * cn->ln);
  if ((rc = visc_be(g, psi2, chi2, b2, hdt))) goto done;
  hyperv_state(g, psi2, chi2, b2, hdt);
  for (long i = 0; i < n; i++) {
    psi->e[i] = psi->e[i] + (dt) * s->n_psi->e[i];
    chi->e[i] = chi->e[i] + (dt) * s->n_chi->e[i];
    b->e[i] = b->e[i] + (dt) * s->n_b->e[i];
  }
  psi->ln = psi->ln + (dt) * s->n_psi->ln;
  chi->ln = chi->ln + (dt) * s->n_chi->ln;
  if ((rc = visc_be(g, psi, chi, b, dt))) goto done;
  hyperv_state(g, psi, chi, b, dt);
  for (long i = 0; i < n; i++) {
    psi->e[i] = (2.0 * psi2->e[i]) - psi->e[i];
    chi->e[i] = (2.0 * chi2->e[i]) - chi->e[i];
    b->e[i] = (2.0 * b2->e[i]) - b->e[i];
  }
  psi->ln = (2.0 * psi2->ln) - psi->ln;
  chi->ln = (2.0 * chi2->ln) - chi->ln;
done:
  og_field_free(psi2); og_field_free(chi2); og_field_free(b2);
  og_field_free(pn); og_field_free(cn); og_field_free(bnn);
  return rc;
}

/* STEP_BOUSSI_AB_CN: mod_march.f90:609-661 */
static int step_ab_cn(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  double dt = g->p.dt;
  int rc;
  s->t += dt;
  s->n += 1;
  og_field *psin = og_field_new(g, OG_FFF), *chin = og_field_new(g, OG_FFF), *bn = og_field_new(g, OG_FFF);
  og_field *psi0 = og_field_new(g, OG_FFF), *chi0 = og_field_new(g, OG_FFF), *b0 = og_field_new(g, OG_FFF);
  long n = og_field_size(g, OG_FFF);
  if ((rc = og_force(g, psi, chi, b, psin, chin, bn))) goto done;
  og_copy(g, psi0, psi);
  og_copy(g, chi0, chi);
  og_copy(g, b0, b);
  for (long i = 0; i < n; i++) {
    psi->e[i] = psi->e[i] + dt * (1.5 * psin->e[i] - 0.5 * s->n_psi->e[i]);
    chi->e[i] = chi->e[i] + dt * (1.5 * chin->e[i] - 0.5 * s->n_chi->e[i]);
    b->e[i] = b->e[i] + dt * (1.5 * bn->e[i] - 0.5 * s->n_b->e[i]);
  }
  psi->ln = psi->ln + dt * (1.5 * psin->ln - 0.5 * s->n_psi->ln);
  chi->ln = chi->ln + dt * (1.5 * chin->ln - 0.5 * s->n_chi->ln);
  og_copy(g, s->n_psi, psin);
  og_copy(g, s->n_chi, chin);
  og_copy(g, s->n_b, bn);
  if ((rc = visc_cn(g, psi, chi, b, psi0, chi0, b0))) goto done;
  hyperv_state(g, psi, chi, b, dt);
done:
  og_field_free(psin); og_field_free(chin); og_field_free(bn);
  og_field_free(psi0); og_field_free(chi0); og_field_free(b0);
  return rc;
}

/* CALC_BOUSSI_ETDFE_SUB / ETDAB_SUB: mod_march.f90:948-1139 */
static void etd_sub(const og_grid *g, const double *e16, const double *nl16, double dt, og_field *rur, og_field *rup,
                    og_field *uz, og_field *b, const og_field *nr_, const og_field *np_, const og_field *nz_,
                    const og_field *nb_, const og_solver *hist) {
#pragma omp parallel for
  for (int kk = 0; kk < g->nxchopdim; kk++) {
    double er[16], ar[16];
    memcpy(er, e16, sizeof(er));
    memcpy(ar, nl16, sizeof(ar));
    double notk0 = 1.0;
    if (kk == 0 && mutation != 8) {
      notk0 = 0.0;
      for (int j = 0; j < 4; j++) { er[0 + 4 * j] = 0; er[1 + 4 * j] = 0; ar[0 + 4 * j] = 0; ar[1 + 4 * j] = 0; }
      er[0] = 1.0; er[1 + 4] = 1.0; ar[0] = dt; ar[1 + 4] = dt;
    }
    for (int mm = 0; mm < g->ntchopdim; mm++) {
      double notm0 = (mm == 0) ? 0.0 : 1.0;
      for (int n = 0; n < g->nrchopdim; n++) {
        size_t id = FIDX(g, n, mm, kk);
        cplx ev[4], av[4];
        if (hist) {
          ev[0] = rur->e[id] - hist->r_pi_r->e[id];
          ev[1] = rup->e[id] - hist->r_pi_p->e[id] * notm0;
          ev[2] = uz->e[id] - hist->pi_z->e[id] * notk0;
          ev[3] = b->e[id];
          av[0] = 1.5 * nr_->e[id] - 0.5 * hist->r_n_r->e[id];
          av[1] = 1.5 * np_->e[id] - 0.5 * hist->r_n_p->e[id];
          av[2] = 1.5 * nz_->e[id] - 0.5 * hist->n_z->e[id];
          av[3] = 1.5 * nb_->e[id] - 0.5 * hist->n_b->e[id];
        } else {
          ev[0] = rur->e[id]; ev[1] = rup->e[id]; ev[2] = uz->e[id]; ev[3] = b->e[id];
          av[0] = nr_->e[id]; av[1] = np_->e[id]; av[2] = nz_->e[id]; av[3] = nb_->e[id];
        }
        cplx cv[4];
        for (int i = 0; i < 4; i++) {
          cplx s1 = 0, s2 = 0;
          for (int j = 0; j < 4; j++) { s1 += er[i + 4 * j] * ev[j]; s2 += ar[i + 4 * j] * av[j]; }
          cv[i] = s1 + s2;
        }
        rur->e[id] = cv[0]; rup->e[id] = cv[1]; uz->e[id] = cv[2]; b->e[id] = cv[3];
      }
    }
  }
}

static void pc2vel3(og_grid *g, const og_field *psi, const og_field *chi, og_field *rur, og_field *rup, og_field *uz) {
  og_chopset(g, 3);
  realloc_space(g, rur, OG_FFF); realloc_space(g, rup, OG_FFF); realloc_space(g, uz, OG_FFF);
  og_pc2vel(g, psi, chi, rur, rup, uz);
  og_chopset(g, -3);
}

/* STEP_BOUSSI_ETDFE_BE: mod_march.f90:664-753 */
static int step_etdfe_be(og_solver *s, og_field *psi, og_field *chi, og_field *b, const double *e16, const double *nl16,
                         double dt, int save) {
  og_grid *g = s->g;
  int rc = 0;
  og_field *rurn = og_field_new(g, OG_FFF), *rupn = og_field_new(g, OG_FFF), *uzn = og_field_new(g, OG_FFF);
  og_field *rorn = og_field_new(g, OG_FFF), *ropn = og_field_new(g, OG_FFF), *ozn = og_field_new(g, OG_FFF);
  og_field *bn = og_field_new(g, OG_FFF), *w = og_field_new(g, OG_FFF);
  long n = og_field_size(g, OG_FFF);
  og_chopset(g, 3);
  rc = og_pc2vel(g, psi, chi, rurn, rupn, uzn);
  if (!rc) rc = og_pc2vor(g, psi, chi, rorn, ropn, ozn);
  og_chopset(g, -3);
  if (rc) goto done;
  if ((rc = og_nonlin(g, b, rurn, rupn, uzn, bn, rorn, ropn, ozn))) goto done;
  etd_hyperv(g, rorn, ropn, ozn, bn, dt);
  pc2vel3(g, psi, chi, rurn, rupn, uzn);
  etd_sub(g, e16, nl16, dt, rurn, rupn, uzn, b, rorn, ropn, ozn, bn, NULL);
  if ((rc = etd_sponge(g, uzn, b, dt))) goto done;
  etd_hyperv(g, rurn, rupn, uzn, b, dt);
  if (save) {
    memcpy(s->n_b->e, bn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->r_n_r->e, rorn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->r_n_p->e, ropn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->n_z->e, ozn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->r_pi_r->e, rurn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->r_pi_p->e, rupn->e, sizeof(cplx) * (size_t)n);
    memcpy(s->pi_z->e, uzn->e, sizeof(cplx) * (size_t)n);
  }
  if ((rc = og_project(g, rurn, rupn, uzn, psi, w, 1, psi->ln))) goto done;
  if ((rc = og_idel2ln(g, w, chi))) goto done;
  if (save) {
    pc2vel3(g, psi, chi, rurn, rupn, uzn);
    for (long i = 0; i < n; i++) {
      s->r_pi_r->e[i] = 0.5 * (s->r_pi_r->e[i] - rurn->e[i]);
      s->r_pi_p->e[i] = 0.5 * (s->r_pi_p->e[i] - rupn->e[i]);
      s->pi_z->e[i] = 0.5 * (s->pi_z->e[i] - uzn->e[i]);
    }
  }
done:
  og_field_free(rurn); og_field_free(rupn); og_field_free(uzn);
  og_field_free(rorn); og_field_free(ropn); og_field_free(ozn);
  og_field_free(bn); og_field_free(w);
  return rc;
}

/* RICH_BOUSSI_ETDFE_BE: mod_march.f90:756-808 */
static int rich_etdfe_be(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  double dt = g->p.dt, hdt = dt * 0.5;
  double he[16], hn[16];
  og_etd_op(g, hdt, he, hn);
  int rc;
  s->t += dt;
  s->n += 1;
  og_field *psi2 = og_field_new(g, OG_FFF), *chi2 = og_field_new(g, OG_FFF), *b2 = og_field_new(g, OG_FFF);
  og_copy(g, psi2, psi);
  og_copy(g, chi2, chi);
  og_copy(g, b2, b);
  long n = og_field_size(g, OG_FFF);
  if ((rc = step_etdfe_be(s, psi2, chi2, b2, he, hn, hdt, 0))) goto done;
  if ((rc = visc_be(g, psi2, chi2, b2, hdt))) goto done;
  if ((rc = step_etdfe_be(s, psi2, chi2, b2, he, hn, hdt, 0))) goto done;
  if ((rc = visc_be(g, psi2, chi2, b2, hdt))) goto done;
  if ((rc = step_etdfe_be(s, psi, chi, b, s->etd_e, s->etd_nl, dt, 1))) goto done;
  if ((rc = visc_be(g, psi, chi, b, dt))) goto done;
  for (long i = 0; i < n; i++) {
    psi->e[i] = (2.0 * psi2->e[i]) - psi->e[i];
    chi->e[i] = (2.0 * chi2->e[i]) - chi->e[i];
    b->e[i] = (2.0 * b2->e[i]) - b->e[i];
  }
  psi->ln = (2.0 * psi2->ln) - psi->ln;
  chi->ln = (2.0 * chi2->ln) - chi->ln;
done:
  og_field_free(psi2); og_field_free(chi2); og_field_free(b2);
  return rc;
}

/* STEP_BOUSSI_ETDAB_CN: mod_march.f90:811-901 */
static int step_etdab_cn(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  double dt = g->p.dt;
  int rc = 0;
  s->t += dt;
  s->n += 1;
  og_field *rurn = og_field_new(g, OG_FFF), *rupn = og_field_new(g, OG_FFF), *uzn = og_field_new(g, OG_FFF);
  og_field *rorn = og_field_new(g, OG_FFF), *ropn = og_field_new(g, OG_FFF), *ozn = og_field_new(g, OG_FFF);
  og_field *bn = og_field_new(g, OG_FFF), *w = og_field_new(g, OG_FFF);
  og_field *psi0 = og_field_new(g, OG_FFF), *chi0 = og_field_new(g, OG_FFF), *b0 = og_field_new(g, OG_FFF);
  long n = og_field_size(g, OG_FFF);
  og_chopset(g, 3);
  rc = og_pc2vel(g, psi, chi, rurn, rupn, uzn);
  if (!rc) rc = og_pc2vor(g, psi, chi, rorn, ropn, ozn);
  og_chopset(g, -3);
  if (rc) goto done;
  if ((rc = og_nonlin(g, b, rurn, rupn, uzn, bn, rorn, ropn, ozn))) goto done;
  etd_hyperv(g, rorn, ropn, ozn, bn, dt);
  og_copy(g, b0, b);
  pc2vel3(g, psi, chi, rurn, rupn, uzn);
  etd_sub(g, s->etd_e, s->etd_nl, dt, rurn, rupn, uzn, b, rorn, ropn, ozn, bn, s);
  memcpy(s->n_b->e, bn->e, sizeof(cplx) * (size_t)n);
  memcpy(s->r_n_r->e, rorn->e, sizeof(cplx) * (size_t)n);
  memcpy(s->r_n_p->e, ropn->e, sizeof(cplx) * (size_t)n);
  memcpy(s->n_z->e, ozn->e, sizeof(cplx) * (size_t)n);
  if ((rc = etd_sponge(g, uzn, b, dt))) goto done;
  etd_hyperv(g, rurn, rupn, uzn, b, dt);
  memcpy(s->r_pi_r->e, rurn->e, sizeof(cplx) * (size_t)n);
  memcpy(s->r_pi_p->e, rupn->e, sizeof(cplx) * (size_t)n);
  memcpy(s->pi_z->e, uzn->e, sizeof(cplx) * (size_t)n);
  og_copy(g, psi0, psi);
  og_copy(g, chi0, chi);
  if ((rc = og_project(g, rurn, rupn, uzn, psi, w, 1, psi->ln))) goto done;
  if ((rc = og_idel2ln(g, w, chi))) goto done;
  pc2vel3(g, psi, chi, rurn, rupn, uzn);
  for (long i = 0; i < n; i++) {
    s->r_pi_r->e[i] = s->r_pi_r->e[i] - rurn->e[i];
    s->r_pi_p->e[i] = s->r_pi_p->e[i] - rupn->e[i];
    s->pi_z->e[i] = s->pi_z->e[i] - uzn->e[i];
  }
  rc = visc_cn(g, psi, chi, b, psi0, chi0, b0);
done:
  og_field_free(rurn); og_field_free(rupn); og_field_free(uzn);
  og_field_free(rorn); og_field_free(ropn); og_field_free(ozn);
  og_field_free(bn); og_field_free(w);
  og_field_free(psi0); og_field_free(chi0); og_field_free(b0);
  return rc;
}

/* ===================================================================== */
/* DIAGNOST's every-IT1 section: mod_march.f90:1797-1961, CALCXY :1967-2047 */
/* ===================================================================== */

/* INTEG: mod_scalar3.f90:312-353 */
static double integ(const og_grid *g, const og_field *f) {
  const double pi = acos(-1.0);
  return 4.0 * pi * g->zlen * g->ell2 * creal(f->e[0]) * NORM(g, 0, 0);
}

/* PRODCT: mod_scalar3.f90:398-468 (PROD(r) over k then m, then the
 * weighted sum over r; m = 0 as a complex product, its real part kept) */
static double prodct(const og_grid *g, const og_field *a, const og_field *b) {
  const double pi = acos(-1.0);
  cplx *prod = (cplx *)calloc((size_t)g->nr, sizeof(cplx));
  for (int kk = 0; kk < g->nxchopdim; kk++)
    for (int mm = 0; mm < g->ntchopdim; mm++)
      for (int r = 0; r < g->nr; r++) {
        const cplx av = a->e[QIDX(g, r, mm, kk)], bv = b->e[QIDX(g, r, mm, kk)];
        if (mm == 0) prod[r] = prod[r] + av * conj(bv);
        else prod[r] = prod[r] + 2.0 * (creal(av) * creal(bv) + cimag(av) * cimag(bv));
      }
  cplx s = 0;
  for (int r = 0; r < g->nr; r++) s = s + prod[r] * g->w[r];
  free(prod);
  return 2 * pi * g->zlen * g->ell2 * creal(s);
}

/* |CALCAT1| (mod_scalar3.f90:286-310) at the current truncation: at k = 0
 * and max - min over k */
static void calcat1_range(const og_grid *g, const og_field *f, double *at0, double *spread) {
  double mx = -1.0, mn = 1e300;
  for (int kk = 0; kk < g->nxchopdim; kk++) {
    const double a = cabs(calcat1(g, f, kk));
    if (kk == 0) *at0 = a;
    mx = fmax(mx, a);
    mn = fmin(mn, a);
  }
  *spread = mx - mn;
}

/* ABSVEC2: mod_march.f90:2050-2099 (a may be w) */
static void absvec2(const og_grid *g, const og_field *u, const og_field *v, const og_field *w, og_field *a) {
  for (int kk = 0; kk < g->ndimx; kk++)
    for (int j = 0; j < g->nth; j++)
      for (int i = 0; i < g->nr; i++) {
        const size_t id = PIDX(g, i, j, kk);
        const double r = g->r[i], f = (1 - g->x[i]) * (1 - g->x[i]);
        const cplx uu = u->e[id], vv = v->e[id], ww = w->e[id];
        const double aaa = ((creal(uu) * creal(uu) + creal(vv) * creal(vv)) / (r * r) + creal(ww) * creal(ww));
        const double bbb = ((cimag(uu) * cimag(uu) + cimag(vv) * cimag(vv)) / (r * r) + cimag(ww) * cimag(ww));
        a->e[id] = aaa / f + I * (bbb / f);
      }
}

/* MULRCOSTH (q = 0), MULRSINTH (1), MULR2 (2): mod_march.f90:2102-2225 */
static void mul_rtheta(const og_grid *g, og_field *a, int q) {
  for (int kk = 0; kk < g->ndimx; kk++)
    for (int j = 0; j < g->nth; j++)
      for (int i = 0; i < g->nr; i++) {
        const size_t id = PIDX(g, i, j, kk);
        const double r = g->r[i];
        double fr, fi;
        if (q == 0) fr = r * cos(g->thr[j]), fi = r * cos(g->thi[j]);
        else if (q == 1) fr = r * sin(g->thr[j]), fi = r * sin(g->thi[j]);
        else fr = fi = r * r;
        a->e[id] = fr * creal(a->e[id]) + I * (fi * cimag(a->e[id]));
      }
}

/* out: VALVOR, VALZMO, VALENE, VALHEL, VALANG, |PSIAT1(k=0)|, spread,
 * |CHIAT1(k=0)|, spread, <X>, <Y>, <R> */
int og_diagnost_it1(og_grid *g, const og_field *psi, const og_field *chi, double *out) {
  const double pi = acos(-1.0);
  og_field *rur = og_field_new(g, OG_FFF), *rup = og_field_new(g, OG_FFF), *uzh = og_field_new(g, OG_FFF);
  og_field *ror = og_field_new(g, OG_FFF), *rop = og_field_new(g, OG_FFF), *ozh = og_field_new(g, OG_FFF);
  long n = og_field_size(g, OG_FFF);
  og_chopset(g, 3);
  /* PC2VEL(..., 'H'): UZ/(1-x)^2 = -DELSQH(chi) (elements only) */
  int rc = og_delsqh(g, chi, uzh);
  for (long i = 0; i < n && !rc; i++) uzh->e[i] = -uzh->e[i];
  /* PC2VOR(..., 'H'): ROR, ROP and OZ/(1-x)^2 = -DELSQH(psi) */
  if (!rc) rc = og_pc2vor(g, psi, chi, ror, rop, ozh);
  if (!rc) rc = og_delsqh(g, psi, ozh);
  for (long i = 0; i < n && !rc; i++) ozh->e[i] = -ozh->e[i];
  if (!rc) {
    out[0] = integ(g, ozh);
    out[1] = integ(g, uzh);
    /* VALANG = -INTEG_P(XXDX(psi, LN = 0)): INTEG_P (mod_march.f90:3649-3682)
     * reads F%E(:NR,1,1) of its FFF argument, the first NR stored
     * coefficients in storage order */
    og_copy(g, rur, psi);
    rur->ln = 0.0;
    rc = og_xxdx(g, rur, rup);
  }
  if (!rc) {
    double dot = 0.0;
    for (int i = 0; i < g->nr; i++) dot = dot + creal(rup->e[i]) / ((1 - g->x[i]) * (1 - g->x[i])) * g->w[i];
    out[4] = -(2.0 * pi * g->zlen * g->ell2 * dot);
    og_copy(g, rup, psi);
    rup->ln = 2 * rup->ln;
    realloc_space(g, rur, OG_FFF);
    rc = og_delsqh(g, rup, rur);
  }
  if (!rc) rc = og_rtran(g, rur, 1);
  if (!rc) {
    og_copy(g, rup, psi);
    rup->ln = 0.0;
    rc = og_rtran(g, rup, 1);
  }
  double ene1 = 0, ene2 = 0;
  if (!rc) {
    ene1 = prodct(g, rur, rup);
    realloc_space(g, rup, OG_FFF);
    rc = og_del2(g, chi, rup);
  }
  if (!rc) rc = og_rtran(g, rup, 1);
  if (!rc) {
    realloc_space(g, rur, OG_FFF);
    rc = og_delsqh(g, chi, rur);
  }
  if (!rc) rc = og_rtran(g, rur, 1);
  if (!rc) {
    ene2 = prodct(g, rur, rup);
    out[2] = -ene1 + ene2;
    realloc_space(g, rur, OG_FFF);
    rc = og_delsqh(g, psi, rur);
  }
  if (!rc) rc = og_rtran(g, rur, 1);
  if (!rc) {
    out[3] = 2 * prodct(g, rur, rup);
    calcat1_range(g, psi, &out[5], &out[6]);
    calcat1_range(g, chi, &out[7], &out[8]);
    realloc_space(g, rur, OG_FFF);
    rc = og_mulxm(g, ozh, rur);
  }
  if (!rc) rc = og_mulxm(g, rur, ozh);
  /* CALCXY */
  if (!rc) rc = og_tofp(g, ror);
  if (!rc) rc = og_tofp(g, rop);
  if (!rc) rc = og_tofp(g, ozh);
  double valens = 0.0, mom[3] = {0, 0, 0};
  if (!rc) {
    absvec2(g, ror, rop, ozh, ozh);
    og_copy(g, ror, ozh);
    rc = og_toff(g, ror);
    valens = integ(g, ror);
  }
  for (int q = 0; q < 3 && !rc; q++) {
    og_copy(g, ror, ozh);
    mul_rtheta(g, ror, q);
    rc = og_toff(g, ror);
    mom[q] = integ(g, ror);
  }
  if (!rc) {
    out[9] = mom[0] / valens;
    out[10] = mom[1] / valens;
    out[11] = sqrt(mom[2] / valens);
  }
  og_chopset(g, -3);
  og_field_free(rur); og_field_free(rup); og_field_free(uzh);
  og_field_free(ror); og_field_free(rop); og_field_free(ozh);
  return rc;
}

/* DIAGNOST per-step checks: mod_march.f90:1726-1791 */
int og_diagnost(const og_grid *g, const og_field *psi, const og_field *chi) {
  (void)chi;
  double chk = 0.0, a = 0.0, b2 = 0.0;
  for (int kk = 0; kk < g->nxchopdim; kk++) {
    for (int n = g->nrchops[0]; n < g->nrchopdim; n++) a = fmax(a, cabs(psi->e[FIDX(g, n, 0, kk)]));
    if (g->ntchopdim >= 2)
      for (int n = g->nrchops[1]; n < g->nrchopdim; n++) b2 = fmax(b2, cabs(psi->e[FIDX(g, n, 1, kk)]));
  }
  chk = a + b2;
  if (chk > 1.0e-12) return E_DEALIAS;
  int kmax = imin(5, g->nxchopdim), mmax = imin(6, g->ntchopdim), nmax = imin(5, g->nrchop);
  for (int kk = 0; kk < kmax; kk++)
    for (int mm = 1; mm < mmax; mm++)
      for (int n = 0; n < nmax; n++)
        if (cabs(psi->e[FIDX(g, n, mm, kk)]) > 100.) return E_DIVERGE;
  return E_OK;
}

/* SOLVER_T%INITIALIZE: mod_march.f90:171-240 (file output omitted) */
int og_initialize(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  if (!g->p.adams && g->p.visc_sw == 2) return E_SOLVER; /* mod_march.f90:190-201 */
  if (g->p.adams) return rich_fe_be(s, psi, chi, b);
  if (g->p.bv0 == 0.0) return E_SOLVER;
  og_etd_op(g, g->p.dt, s->etd_e, s->etd_nl);
  return rich_etdfe_be(s, psi, chi, b);
}

/* The ETD operators of SOLVER_T%INITIALIZE without the Richardson start
 * (used to time steps from a given state). */
int og_solver_prepare(og_solver *s) {
  og_grid *g = s->g;
  if (g->p.adams) return E_OK;
  if (g->p.bv0 == 0.0) return E_SOLVER;
  og_etd_op(g, g->p.dt, s->etd_e, s->etd_nl);
  return E_OK;
}

/* TIME_STEPPING's REMOVE and HYP3ADJ cadence (RMV%SW, RMV%INT, VISC%ADJSW,
 * VISC%ADJINT); rmv_sw = 2 also turns on the ETD steps' axial sponge
 * (SPNGINIT, made by INITIALIZE: set before og_initialize) */
void og_set_march_controls(og_solver *s, int rmv_sw, int rmv_int, int adj_sw, int adj_int) {
  s->rmv_sw = rmv_sw;
  s->rmv_int = rmv_int;
  s->adj_sw = adj_sw;
  s->adj_int = adj_int;
  og_set_sponge(s->g, rmv_sw);
}

/* SOLVER_T%TIME_STEPPING: mod_march.f90:243-307 (step, per-step checks,
 * REMOVE with the Richardson restart every RMV%INT steps, HYP3ADJ and
 * HYP3SET every ADJINT steps; file output omitted) */
int og_time_step(og_solver *s, og_field *psi, og_field *chi, og_field *b) {
  og_grid *g = s->g;
  int rc = g->p.adams ? step_ab_cn(s, psi, chi, b) : step_etdab_cn(s, psi, chi, b);
  if (rc) return rc;
  if ((rc = og_diagnost(g, psi, chi))) return rc;
  if (s->rmv_sw != 0 && s->rmv_int > 0 && s->n % s->rmv_int == 0) {
    if ((rc = og_remove(g, psi, chi, b))) return rc;
    rc = g->p.adams ? rich_fe_be(s, psi, chi, b) : rich_etdfe_be(s, psi, chi, b);
    if (rc) return rc;
    if ((rc = og_diagnost(g, psi, chi))) return rc;
  }
  if (s->adj_sw != 0 && s->adj_int > 0 && s->n % s->adj_int == 0) rc = og_hyp3adj(g, psi, chi, b);
  return rc;
}

/* QVORTEX (single centred Gaussian, H = 0): mod_layout.f90:449-610 with
 * LAY (:64-118) and GAUS (:640-673).  omega_z = 2 q exp(-r^2). */
int og_qvortex(og_grid *g, double q, og_field *psi, og_field *chi) {
  og_field *a = og_field_new(g, OG_PPP);
  fzero(g, a);
  for (int z = 0; z < g->ndimx; z++)
    for (int j = 0; j < g->nth; j++)
      for (int r = 0; r < g->nr; r++) {
        double rr = g->r[r];
        double f = pow(1 - g->x[r], 2);
        double gv = (q * 2.0) / pow(1.0, 2.0) * exp(-1.0 * ((pow(rr * cos(g->thr[j]), 2.0) + pow(rr * sin(g->thr[j]), 2.0)) / pow(1.0, 2.0)));
        double gi = (q * 2.0) / pow(1.0, 2.0) * exp(-1.0 * ((pow(rr * cos(g->thi[j]), 2.0) + pow(rr * sin(g->thi[j]), 2.0)) / pow(1.0, 2.0)));
        a->e[PIDX(g, r, j, z)] += (gv + I * gi) / f;
      }
  og_chopdo(g, a);
  int rc = og_toff(g, a);
  long n = og_field_size(g, OG_FFF);
  for (long i = 0; i < n; i++) a->e[i] = -a->e[i];
  a->ln = 0.0;
  og_field *bb = og_field_new(g, OG_FFF);
  if (!rc) rc = og_idelsqh(g, a, bb);
  og_copy(g, psi, bb);
  realloc_space(g, chi, OG_FFF);
  memset(chi->e, 0, sizeof(cplx) * (size_t)n);
  og_field_free(a);
  og_field_free(bb);
  return rc;
}
