// Host-side set-up tables for libmlegs_hip.so (built once per context).
//
// Gauss-Legendre nodes/weights (GAUSS_LEGENDRE, src/mod_lin_legendre.f90:79-131),
// normalisation logs (LEG_LOG_NORMT :306-377, LEG_NORMT :205-278), the
// orthonormal associated-Legendre table (LEG_TBL3 :526-644) and the banded
// operator coefficients (BAND_LOGLEG_XM :910-982, _XXDX :1683-1758,
// _RAT_DEL2H :1986-2092).  The reference carries the Legendre recurrence in
// 50-digit FM arithmetic; here it runs in 80-bit long double, which keeps
// the 1-ulp-of-FP64 accuracy the table needs and the exponent range that
// P_n^m ~ (1-x^2)^{m/2} needs at m <= 85 near |x| = 1.
#include "mlegs_internal.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace mlg {

void gauss_legendre(int n, double *x, double *w) {
  const double pi = std::acos(-1.0);
  int mh = (int)std::ceil((n + 1) / 2.0);
  for (int i = 1; i <= mh; i++) {
    double z = std::cos(pi * (i - 0.25) / (n + 0.5));
    double z1, pp;
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
    } while (std::fabs(z - z1) > 1e-15);
    x[i - 1] = -z;
    x[n - i] = z;
    w[i - 1] = 2.0 / ((1 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

void leg_log_norm(int ndim, int m, double *out) {
  double wk = std::log(0.5);
  for (int mm = 1; mm <= m; mm++)
    wk = wk + std::log(2 * mm + 1.0) - std::log((double)(2 * mm) * ((2 * mm - 1.0) * (2 * mm - 1.0)));
  out[0] = wk;
  for (int nn = 2; nn <= ndim; nn++) {
    int n = m + nn - 1;
    out[nn - 1] = out[nn - 2] + std::log(2 * n + 1.0) - std::log(2 * n - 1.0) + std::log(1.0 * (n - m)) -
                  std::log(1.0 * (n + m));
  }
  for (int i = 0; i < ndim; i++) out[i] = 0.5 * out[i];
}

void leg_norm(int ndim, int m, double *out) {
  double wk = 0.5;
  for (int mm = 1; mm <= m; mm++) wk = wk * (2 * mm + 1.0) / ((double)(2 * mm) * ((2 * mm - 1.0) * (2 * mm - 1.0)));
  out[0] = wk * std::pow(100.0, m);
  for (int nn = 2; nn <= ndim; nn++) {
    int n = m + nn - 1;
    out[nn - 1] = out[nn - 2] * (2 * n + 1.0) / (2 * n - 1.0) * (n - m) / (n + m);
  }
  for (int i = 0; i < ndim; i++) out[i] = std::sqrt(out[i]) * std::pow(10.0, -m);
}

static double logfact(int n) { return std::lgamma(2 * n + 1.0) - n * std::log(2.0) - std::lgamma(n + 1.0); }

// out[xx + nx*(nn-1)] = Pbar_{m+nn-1}^m(x[xx]), nn = 1..ne
void leg_tbl(const double *x, int nx, int ne, int m, const double *lognorm, double *out) {
  std::vector<long double> q((size_t)ne);
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
      q[nn - 1] = 1.0L / (n - m) *
                  (q[nn - 2] * xl - q[nn - 3] * (long double)(n + m - 1) / (long double)(2 * n - 1) / (long double)(2 * n - 3));
    }
    for (int nn = 1; nn <= ne; nn++) {
      long double e = (long double)lognorm[nn - 1] + (long double)logfact(m + nn - 1);
      double d = (double)(q[nn - 1] * expl(e));
      if (std::fabs(d) < DBL_MIN) d = 0.0;
      out[xx + (size_t)nx * (nn - 1)] = d;
    }
  }
}

// Band rows: b[row*nb + col] with col 0..nb-1, centre nb/2; row n <-> degree m+n.
void band_xm(int ni, int m, const double *ln, double *b) {
  std::memset(b, 0, sizeof(double) * 3 * (size_t)ni);
  for (int nn = 1; nn <= ni; nn++) {
    int n = m + nn - 1;
    double *r = b + (size_t)3 * (nn - 1);
    if (nn - 1 >= 1) r[0] = -((double)(n - m)) / (2 * n - 1.0) * std::exp(ln[nn - 2] - ln[nn - 1]);
    r[1] = 1.0;
    if (nn + 1 <= ni) r[2] = -(n + m + 1.0) / (2 * n + 3.0) * std::exp(ln[nn] - ln[nn - 1]);
  }
}

void band_xxdx(int ni, int m, const double *ln, double *b) {
  std::memset(b, 0, sizeof(double) * 3 * (size_t)ni);
  for (int nn = 1; nn <= ni; nn++) {
    int n = m + nn - 1;
    double *r = b + (size_t)3 * (nn - 1);
    if (nn - 1 >= 1) r[0] = -(n - 1.0) * (n - m) / (2 * n - 1.0) * std::exp(ln[nn - 2] - ln[nn - 1]);
    if (nn + 1 <= ni) r[2] = (n + 2.0) * (n + m + 1.0) / (2 * n + 3.0) * std::exp(ln[nn] - ln[nn - 1]);
  }
}

// L^2 times the horizontal Laplacian, (1-x)^-2 del_h^2, on the mapped
// Legendre functions (Matsushima & Marcus 1997): with n the degree, the image
// of P_n^m spans degrees n-2 .. n+2, with coefficients
//   d-2: -(n-m-1)(n-m)(n-2)(n-1) / ((2n-3)(2n-1))
//   d-1:  2n(n-m)(n-1) / (2n-1)
//   d 0: -2n(n+1)(3n^2+3n-m^2-2) / ((2n-1)(2n+3))
//   d+1:  2(n+1)(n+m+1)(n+2) / (2n+3)
//   d+2: -(n+m+1)(n+m+2)(n+2)(n+3) / ((2n+3)(2n+5))
// for the unnormalised P_n^m, divided by L^2; the off-diagonals are rescaled
// to the normalised basis by the ratio of the log-norms.  The numerators are
// integers below 2^53 (exact in FP64); the divisions are made one factor at a
// time.
void band_del2h(int ni, int m, double ell, const double *ln, double *b) {
  const double l2 = std::pow(ell, 2.0);
  std::memset(b, 0, sizeof(double) * 5 * (size_t)ni);
  for (int row = 0; row < ni; row++) {
    const long n = m + row;
    double *r = b + (size_t)5 * row;
    const double up = 2.0 * n + 3.0;  // 2n+3
    const double lo = 2.0 * n - 1.0;  // 2n-1
    if (row >= 2) {
      const double num = (double)((n - m - 1) * (n - m) * (n - 2) * (n - 1));
      r[0] = -num / (2.0 * n - 3.0) / lo / l2 * std::exp(ln[row - 2] - ln[row]);
    }
    if (row >= 1) r[1] = (double)(2 * n * (n - m) * (n - 1)) / lo / l2 * std::exp(ln[row - 1] - ln[row]);
    r[2] = (double)(-2 * n * (n + 1) * (3 * n * n + 3 * n - (long)m * m - 2)) / lo / up / l2;
    if (row + 1 < ni) r[3] = (double)(2 * (n + 1) * (n + m + 1) * (n + 2)) / up / l2 * std::exp(ln[row + 1] - ln[row]);
    if (row + 2 < ni) {
      const double num = (double)((n + m + 1) * (n + m + 2) * (n + 2) * (n + 3));
      r[4] = -num / up / (2.0 * n + 5.0) / l2 * std::exp(ln[row + 2] - ln[row]);
    }
  }
}

// HYP3MASK / HYP3SET: src/mod_march.f90:3320-3480
static void hyp3mask(const double *kv, int n, double ratio, double steep, double *mask) {
  double kmax = -1e300;
  for (int i = 0; i < n; i++) kmax = kv[i] > kmax ? kv[i] : kmax;
  if (kmax < 1.0e-12) kmax = 1.0;
  double ta = std::tanh(steep);
  for (int i = 0; i < n; i++) {
    double kn = kv[i] / kmax;
    if (kn < ratio) {
      mask[i] = 0.0;
    } else {
      double xr = (kn - ratio) / (1.0 - ratio);
      double yr = steep * (2.0 * xr - 1.0);
      mask[i] = (std::tanh(yr) + ta) / (2.0 * ta);
    }
  }
}

// hr[nn + nrc*mm], ht[mm], hx[kk]; nu_r, nu_t, nu_x per direction
void hyp3set(const Grid &g, double nu_r, double nu_t, double nu_x, std::vector<double> &hr, std::vector<double> &ht,
             std::vector<double> &hx) {
  int nrc = g.p.nrchop, ntd = g.ntchopdim, nxd = g.nxchopdim, ph = g.p.visc_p / 2;
  hr.assign((size_t)nrc * ntd, 0.0);
  ht.assign((size_t)ntd, 0.0);
  hx.assign((size_t)nxd, 0.0);
  std::vector<double> mr((size_t)nrc * ntd, 1.0), mt(ntd), mx(nxd), kv((size_t)std::max(nrc, nxd) + ntd);
  for (int mm = 0; mm < ntd; mm++) {
    int nn = std::max(std::min(nrc, nrc - g.m[mm]), 0);
    for (int i = 0; i < nn; i++) kv[i] = (double)i;
    if (nn > 0) hyp3mask(kv.data(), nn, 0.5, 2.0, mr.data() + (size_t)nrc * mm);
  }
  for (int mm = 0; mm < ntd; mm++) kv[mm] = (double)g.m[mm];
  hyp3mask(kv.data(), ntd, 2.0 / 3.0, 2.0, mt.data());
  for (int kk = 0; kk < nxd; kk++) kv[kk] = std::fabs(g.ak[(size_t)ntd * kk]);
  hyp3mask(kv.data(), nxd, 2.0 / 3.0, 2.0, mx.data());
  double k2r = 1.5 * 1.5 * (double)((nrc - 1) * nrc) / g.ell2;
  double k2t = 0, k2x = 0;
  for (int mm = 0; mm < ntd; mm++) k2t = std::fmax(k2t, (double)(g.m[mm] * g.m[mm]));
  for (int kk = 0; kk < nxd; kk++) k2x = std::fmax(k2x, g.ak[(size_t)ntd * kk] * g.ak[(size_t)ntd * kk]);
  for (int mm = 0; mm < ntd; mm++) {
    ht[mm] = std::exp(-nu_t * mt[mm] * std::pow((double)(g.m[mm] * g.m[mm]) / k2t, ph));
    for (int nn = 0; nn < nrc; nn++)
      hr[nn + (size_t)nrc * mm] =
          std::exp(-nu_r * mr[nn + (size_t)nrc * mm] * std::pow(1.5 * 1.5 * (double)(nn * (nn + 1)) / g.ell2 / k2r, ph));
  }
  for (int kk = 0; kk < nxd; kk++) {
    double a = g.ak[(size_t)ntd * kk];
    hx[kk] = std::exp(-nu_x * mx[kk] * std::pow(a * a / k2x, ph));
  }
}

// e^{sign*2*pi*i*j/n} for j < n, from long double
void twiddles(int n, int sign, double *re, double *im) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; j++) {
    long double a = sign * 2.0L * pi * (long double)j / (long double)n;
    re[j] = (double)cosl(a);
    im[j] = (double)sinl(a);
  }
}

// Builds every host-side table of the grid: LEGINIT, src/mod_fft.f90:180-367
int build_grid(const mlg_params &p, Grid &g) {
  g.p = p;
  if (g.p.minc <= 0) g.p.minc = 1;
  g.nr = p.nr;
  g.nrh = p.nr / 2;
  g.nth = p.nth;
  g.nx = p.nx;
  g.ntchop = p.ntchop;
  g.nxchop = p.nxchop;
  g.ell = p.ell;
  g.ell2 = std::pow(p.ell, 2.0);
  g.zlen = p.zlen;
  g.ndimr = g.nr + 3;
  g.ndimth = g.nth + 1;
  if (g.ndimth == 2) g.ndimth = 1;
  g.ndimx = g.nx;
  g.nrchopdim = std::min(p.nrchop + 6, g.nr);
  g.ntchopdim = std::min(p.ntchop, g.nth);
  g.nxchopdim = 2 * p.nxchop - 1;
  g.nxchoph = g.nx - p.nxchop + 2;
  if (p.nxchop == 1) g.nxchoph = 1;
  g.nrchop = p.nrchop;
  g.npf = g.nrchopdim + 1;
  g.nnorm = g.nrchopdim + 14;
  g.nband = g.nrchopdim + 2;
  g.m.resize(g.ntchopdim);
  for (int mm = 0; mm < g.ntchopdim; mm++) g.m[mm] = g.p.minc * mm;
  const double pi = std::acos(-1.0);
  g.ak.assign((size_t)g.ntchopdim * g.nxchopdim, 0.0);
  for (int mm = 0; mm < g.ntchopdim; mm++)
    for (int kk = 1; kk <= g.nxchopdim; kk++) {
      int kv = kk - 1 + p.mklink * mm;
      if (kk > p.nxchop) kv = -(g.nxchopdim - kk + 1);
      g.ak[mm + (size_t)g.ntchopdim * (kk - 1)] = 2 * pi / g.zlen * kv;
    }
  g.x.resize(g.nr);
  g.w.resize(g.nr);
  g.r.resize(g.nr);
  g.lnx.resize(g.nr);
  gauss_legendre(g.nr, g.x.data(), g.w.data());
  for (int i = 0; i < g.nr; i++) {
    g.r[i] = g.ell * std::sqrt((1 + g.x[i]) / (1 - g.x[i]));
    g.lnx[i] = -std::log(1 - g.x[i]);
  }
  g.norm.assign((size_t)g.nnorm * g.ntchopdim, 0.0);
  g.lognorm.assign((size_t)g.nnorm * g.ntchopdim, 0.0);
  for (int mm = 0; mm < g.ntchopdim; mm++) {
    leg_norm(g.nnorm, g.m[mm], &g.norm[(size_t)g.nnorm * mm]);
    leg_log_norm(g.nnorm, g.m[mm], &g.lognorm[(size_t)g.nnorm * mm]);
  }
  g.pf.assign((size_t)g.nrh * g.npf * g.ntchopdim, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int mm = 0; mm < g.ntchopdim; mm++)
    leg_tbl(g.x.data(), g.nrh, g.npf, g.m[mm], &g.lognorm[(size_t)g.nnorm * mm], &g.pf[(size_t)g.nrh * g.npf * mm]);
  g.at0.resize(g.nrchopdim);
  g.at1.resize(g.nrchopdim);
  double xm1 = -1.0, xp1 = 1.0;
  leg_tbl(&xm1, 1, g.nrchopdim, 0, &g.lognorm[0], g.at0.data());
  leg_tbl(&xp1, 1, g.nrchopdim, 0, &g.lognorm[0], g.at1.data());
  // band operator rows per m, for rows 0..nband-1
  g.bxm.assign((size_t)g.ntchopdim * g.nband * 3, 0.0);
  g.bxxdx.assign((size_t)g.ntchopdim * g.nband * 3, 0.0);
  g.bdel2h.assign((size_t)g.ntchopdim * g.nband * 5, 0.0);
  for (int mm = 0; mm < g.ntchopdim; mm++) {
    const double *ln = &g.lognorm[(size_t)g.nnorm * mm];
    band_xm(g.nband, g.m[mm], ln, &g.bxm[(size_t)mm * g.nband * 3]);
    band_xxdx(g.nband, g.m[mm], ln, &g.bxxdx[(size_t)mm * g.nband * 3]);
    band_del2h(g.nband, g.m[mm], g.ell, ln, &g.bdel2h[(size_t)mm * g.nband * 5]);
  }
  // PROJECT tables (src/mod_legops.f90:2476-2494): V, D, T with rows n and
  // columns over the first NRH nodes.  Rows do not depend on the truncation
  // (PFD(:,c) uses PF columns c-1 and c+1 only), so they are built once for
  // the largest row count.
  int nrow = g.nrchopdim, nrh = g.nrh;
  g.vdt.assign((size_t)3 * g.ntchopdim * nrow * nrh, 0.0);
  std::vector<double> pfd((size_t)nrh * (nrow + 1)), xb((size_t)3 * (nrow + 2));
  for (int mm = 0; mm < g.ntchopdim; mm++) {
    int mv = g.m[mm];
    int nb = nrow + 1;
    band_xxdx(nb, mv, &g.lognorm[(size_t)g.nnorm * mm], xb.data());
    const double *pfm = &g.pf[(size_t)nrh * g.npf * mm];
    for (int i = 0; i < nrh; i++)
      for (int c = 0; c < nb; c++) {
        // vector-times-band, terms in band-column order (sub, diag, super)
        double v = 0.0;
        if (c + 1 < nb) v += xb[(size_t)3 * (c + 1) + 0] * pfm[i + (size_t)nrh * (c + 1)];
        v += xb[(size_t)3 * c + 1] * pfm[i + (size_t)nrh * c];
        if (c - 1 >= 0) v += xb[(size_t)3 * (c - 1) + 2] * pfm[i + (size_t)nrh * (c - 1)];
        pfd[i + (size_t)nrh * c] = v;
      }
    for (int ii = 0; ii < nrow; ii++) {
      int n = std::max(1, mv + ii);
      for (int j = 0; j < nrh; j++) {
        double fff = (1 - g.x[j] * g.x[j]) / g.w[j];
        double den = (double)(n * (n + 1)) * fff;
        size_t base = ((size_t)mm * nrow + ii) * nrh + j;
        g.vdt[base] = pfm[j + (size_t)nrh * ii] / den;
        g.vdt[(size_t)g.ntchopdim * nrow * nrh + base] = pfd[j + (size_t)nrh * ii] / den;
        g.vdt[(size_t)2 * g.ntchopdim * nrow * nrh + base] = pfm[j + (size_t)nrh * ii] * g.w[j];
      }
    }
  }
  if (p.visc_sw == 3 || p.hyper_cn) {
    hyp3set(g, p.nup, p.nup, p.nup, g.hv_r, g.hv_t, g.hv_x);
    hyp3set(g, p.kappap, p.kappap, p.kappap, g.hd_r, g.hd_t, g.hd_x);
  }
  return 0;
}

}  // namespace mlg
