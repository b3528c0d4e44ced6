// Vorticity removal (REMOVE, RINGCUT, SMOOTH), the axial sponge layer
// (SPNGINIT, SPONGE, CALC_BOUSSI_ETD_SPONGE, TOPHAT) and the adaptive
// hyperviscosity (ENEMON, SPECURV, SPESLOP, PDAJUST, HYP3ADJ + HYP3SET) of
// src/mod_march.f90, with the cadence that SOLVER_T%TIME_STEPPING
// (:243-307) gives them.
//
// The per-element work runs on the GPU: the sponge's z-plane damping in
// PFP space (k_sponge_planes) and RINGCUT's zeroing and five weighted
// smoothing passes, one block per (m, k) column staged in LDS
// (k_ringcut).  The spectra's sums, the controllers and the tables
// (TOPHAT_X, HYP3SET) are host work on a few hundred numbers.
#include <cmath>
#include <cstring>
#include <vector>

#include "mlegs_internal.h"

using namespace mlg;

#define RET(x)                 \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

namespace mlg {

// ---------------------------------------------------------------- sponge
// SPNGINIT (src/mod_march.f90:2292-2323): TOPHAT_X(z) = TOPHAT(z - ZLEN/2,
// ZLEN RATIO_L / 2, ZLEN RATIO_S) with RATIO_L = 0.86, RATIO_S = 0.04
// (REAL(8): the reference compiles default reals as 8-byte), TAU = 10 DT.
// TOPHAT (:3704-3712) = (tanh((x + L)/S) - tanh((x - L)/S)) / 2.
static int sponge_init(mlg_ctx *c, int rmv_sw) {
  c->rmv_sw = rmv_sw;
  c->tophat.clear();
  if (rmv_sw != 2) return 0;
  const Grid &G = c->g;
  const double ratio_l = 0.86, ratio_s = 0.04;
  const double hlfz = G.zlen / 2.0, hw = G.zlen * ratio_l / 2.0, st = G.zlen * ratio_s;
  c->sp_tau = 10 * G.p.dt;
  c->tophat.assign((size_t)G.ndimx, 0.0);
  for (int k = 0; k < G.nx; k++) {
    const double x = G.zlen * k / G.nx - hlfz;  // TFM%Z(k) = ZLEN (k-1) / NX, 1-based
    c->tophat[k] = 0.5 * (std::tanh((x + hw) / st) - std::tanh((x - hw) / st));
  }
  if (!c->d_sponge) MLG_CHECK_HIP(hipMalloc((void **)&c->d_sponge, sizeof(double) * 3 * (size_t)G.ndimx));
  // damping of one application over dt and over dt / 2 (the Richardson
  // start's half steps)
  for (int slot = 0; slot < 2; slot++) {
    const double dt = slot ? G.p.dt * 0.5 : G.p.dt;
    std::vector<double> d((size_t)G.ndimx);
    for (int k = 0; k < G.ndimx; k++) d[k] = std::exp(-(1.0 - c->tophat[k]) / c->sp_tau * dt);
    MLG_CHECK_HIP(hipMemcpy(c->d_sponge + (size_t)slot * G.ndimx, d.data(), sizeof(double) * G.ndimx,
                            hipMemcpyHostToDevice));
  }
  return 0;
}

// the device damping of one application over dt: exp(-(1 - TOPHAT_X) / TAU DT)
static int sponge_table(mlg_ctx *c, double dt, const double **out) {
  const Grid &G = c->g;
  if (dt == G.p.dt) {
    *out = c->d_sponge;
    return 0;
  }
  if (dt == G.p.dt * 0.5) {
    *out = c->d_sponge + G.ndimx;
    return 0;
  }
  std::vector<double> d((size_t)G.ndimx);
  for (int k = 0; k < G.ndimx; k++) d[k] = std::exp(-(1.0 - c->tophat[k]) / c->sp_tau * dt);
  double *slot = c->d_sponge + 2 * (size_t)G.ndimx;
  MLG_CHECK_HIP(hipMemcpyAsync(slot, d.data(), sizeof(double) * G.ndimx, hipMemcpyHostToDevice, c->stream));
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  *out = slot;
  return 0;
}

// PFP field (r, m < NTCHOPDIM in the NDIMTH slots, z): every z plane times
// damp[z], then m = 0 made real
__global__ void __launch_bounds__(256) k_sponge_planes(KGrid g, double2 *a, const double *__restrict__ damp) {
  const long n = (long)g.ndimr * g.ntchopdim * g.ndimx;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const int r = (int)(i % g.ndimr);
    const long col = i / g.ndimr;
    const int mm = (int)(col % g.ntchopdim), z = (int)(col / g.ntchopdim);
    const size_t at = (size_t)r + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z);
    const double d = damp[z];
    double2 v = a[at];
    v.x = v.x * d;
    v.y = v.y * d;
    if (mm == 0) v.y = 0.0;
    a[at] = v;
  }
}

static int launch_sponge_planes(mlg_ctx *c, double2 *a, const double *damp) {
  KGrid g = kgrid(c);
  const long n = (long)g.ndimr * g.ntchopdim * g.ndimx;
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sponge_planes, dim3(blocks), dim3(256), 0, c->stream, g, a, damp);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// RTRAN(f, 1); VERFFT(f, 1) for each field; damping; TOFF each
static int sponge_fields(mlg_ctx *c, mlg_field *const *f, int nf, double dt) {
  const double *damp = nullptr;
  RET(sponge_table(c, dt, &damp));
  for (int i = 0; i < nf; i++) {
    RET(mlg_rtran(f[i], 1));
    RET(mlg_verfft(f[i], 1));
  }
  for (int i = 0; i < nf; i++) RET(launch_sponge_planes(c, f[i]->e, damp));
  for (int i = 0; i < nf; i++) RET(mlg_toff(f[i]));
  return 0;
}

// CALC_BOUSSI_ETD_SPONGE (src/mod_march.f90:1436-1469)
int etd_sponge(mlg_ctx *c, mlg_field *uz, mlg_field *b, double dt) {
  if (c->rmv_sw != 2) return 0;
  if (uz->space != MLG_FFF) return 2;  // ERR_FLAGS%TRANSFORM
  mlg_field *const f[2] = {uz, b};
  return sponge_fields(c, f, 2, dt);
}

// ---------------------------------------------------------------- RINGCUT
// One block per (m, k) column of a PFF field; rows ns-1 .. NR-1 (0-based)
// staged in LDS.  Rows ns0-1 .. NR-1 are zeroed except at (m, k) = (0, 0)
// (k > 0 only when NX > 1); the region is weighted by w (divxm2: w /
// (1-x)^2), smoothed 5 times and unweighted.  SMOOTH (src/mod_march.f90:
// 3562-3600) is a sequential scatter; here each output sums the same terms
// in the same order the scatter adds them (the start value, then the
// contributions of inputs j-1, j, j+1), so the results agree bit for bit.
__global__ void __launch_bounds__(256) k_ringcut(KGrid g, double2 *a, int ns0, int ns, int divxm2,
                                                 const double *__restrict__ w, const double *__restrict__ x) {
  extern __shared__ double2 lds[];
  const int ni = g.nr - ns + 1, mm = blockIdx.x % g.ntchopdim, kk = blockIdx.x / g.ntchopdim;
  double2 *va = lds, *vt = lds + ni;
  double2 *col = a + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk) + (ns - 1);
  const bool keep = (mm == 0) && (kk == 0 || g.nx <= 1);
  for (int i = threadIdx.x; i < ni; i += blockDim.x) {
    double2 v = col[i];
    const int r = ns - 1 + i;
    if (!keep && r >= ns0 - 1) v = make_double2(0.0, 0.0);
    const double om = 1.0 - x[r];
    const double wf = divxm2 ? w[r] / (om * om) : w[r];
    v.x = v.x * wf;
    v.y = v.y * wf;
    va[i] = v;
  }
  __syncthreads();
  for (int pass = 0; pass < 5; pass++) {
    for (int j = threadIdx.x; j < ni; j += blockDim.x) {
      double2 t = make_double2(0.0, 0.0);
      if (j == 0) t = va[0];
      if (j >= ni - 3) {
        const double fae = j == ni - 3 ? 0.2 : (j == ni - 2 ? 0.3 : 0.5);
        t.x = t.x + va[ni - 1].x * fae;
        t.y = t.y + va[ni - 1].y * fae;
      }
      // inputs q = j-1, j, j+1 in 1..ni-2 (the scatter's I = q+1 in 2..NI-1),
      // F = (1 + (NI - I) / (NI - 1)) / 2
      for (int q = j - 1; q <= j + 1; q++) {
        if (q < 1 || q > ni - 2) continue;
        const double f = (1 + ((ni - (q + 1)) / (ni - 1.0))) * 0.5;
        const double2 v = va[q];
        if (q == j) {
          t.x = t.x + v.x * f;
          t.y = t.y + v.y * f;
        } else {
          t.x = t.x + v.x * (1 - f) / 2;
          t.y = t.y + v.y * (1 - f) / 2;
        }
      }
      vt[j] = t;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ni; j += blockDim.x) va[j] = vt[j];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < ni; i += blockDim.x) {
    const int r = ns - 1 + i;
    const double om = 1.0 - x[r];
    const double wf = divxm2 ? w[r] / (om * om) : w[r];
    double2 v = va[i];
    v.x = v.x / wf;
    v.y = v.y / wf;
    col[i] = v;
  }
}

static int ringcut(mlg_field *wk, int ns0, int ns, int divxm2) {
  mlg_ctx *c = wk->ctx;
  const Grid &G = c->g;
  const int ni = G.nr - ns + 1;
  if (wk->space != MLG_PFF) return 2;
  if (ns < 1 || ns0 < 1 || ni < 3) return MLG_E_ARG;  // SMOOTH: SIZE TOO SMALL
  KGrid g = kgrid(c);
  const size_t lds = sizeof(double2) * 2 * (size_t)ni;
  if (lds > 64 * 1024) return MLG_E_ARG;
  hipLaunchKernelGGL(k_ringcut, dim3(G.ntchopdim * G.nxchopdim), dim3(256), lds, c->stream, g, wk->e, ns0, ns, divxm2,
                     c->d.w, c->d.x);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- spectra
// the per-(m, k) products of PRODCT_MK_HALFK (l2 = 0: Re(a conj b); 1:
// L2NORM_MK_HALFK's |a|^2 / (1-x)^2) folded to the NXCHOP non-negative k
// (src/mod_scalar3.f90:470-554)
static int halfk(mlg_ctx *c, const double2 *a, const double2 *b, int l2, std::vector<double> &out) {
  const Grid &G = c->g;
  const int nt = G.ntchop, nxc = G.nxchop, nxd = G.nxchopdim;
  const size_t na = (size_t)nt * nxd;
  double *dall = c->d.red;
  RET(launch_energy_terms(c, a, b, l2, dall));
  std::vector<double> h(na);
  MLG_CHECK_HIP(hipMemcpyAsync(h.data(), dall, sizeof(double) * na, hipMemcpyDeviceToHost, c->stream));
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  out.assign((size_t)nt * nxc, 0.0);
  for (int mm = 0; mm < nt; mm++) out[mm] = h[mm];
  for (int kk = 2; kk <= nxc; kk++) {
    const int kc = nxd + 2 - kk;
    for (int mm = 0; mm < nt; mm++) out[mm + (size_t)nt * (kk - 1)] = h[mm + (size_t)nt * (kc - 1)] + h[mm + (size_t)nt * (kk - 1)];
  }
  return 0;
}

// SUM(A, DIM=2) and SUM(A, DIM=1) of an nt x nxc column-major array, in
// order; rows mrows of the first
static void sums(const std::vector<double> &a, int nt, int nxc, int mrows, double *pm, double *pk) {
  for (int mm = 0; mm < mrows; mm++) {
    double s = 0.0;
    for (int kk = 0; kk < nxc; kk++) s = s + a[mm + (size_t)nt * kk];
    pm[mm] = s;
  }
  for (int kk = 0; kk < nxc; kk++) {
    double s = 0.0;
    for (int mm = 0; mm < nt; mm++) s = s + a[mm + (size_t)nt * kk];
    pk[kk] = s;
  }
}

// ENEMON (src/mod_march.f90:2581-2617) at the current truncation
static int enemon(const mlg_field *psi, const mlg_field *chi, double *em, double *ek) {
  mlg_ctx *c = psi->ctx;
  const Grid &G = c->g;
  const int nt = G.ntchop, nxc = G.nxchop;
  Temps T(c);
  mlg_field *w1 = T.get(), *w2 = T.get();
  RET(T.ready());
  std::vector<double> mk;
  std::vector<double> am(nt), ak(nxc);
  RET(copy_field(w1, psi));
  RET(launch_set_ln(c, w1->ln, psi->ln, 2.0, 0.0));  // W1%LN = 2 PSI%LN
  RET(mlg_delsqh(w1, w2));
  RET(mlg_rtran(w2, 1));
  RET(copy_field(w1, psi));
  RET(launch_set_ln(c, w1->ln, nullptr, 0.0, 0.0));
  RET(mlg_rtran(w1, 1));
  RET(halfk(c, w2->e, w1->e, 0, mk));
  sums(mk, nt, nxc, nt, em, ek);
  RET(mlg_del2(chi, w1));
  RET(mlg_rtran(w1, 1));
  RET(mlg_delsqh(chi, w2));
  RET(mlg_rtran(w2, 1));
  RET(halfk(c, w2->e, w1->e, 0, mk));
  sums(mk, nt, nxc, nt, am.data(), ak.data());
  for (int mm = 0; mm < nt; mm++) em[mm] = -em[mm] + am[mm];
  for (int kk = 0; kk < nxc; kk++) ek[kk] = -ek[kk] + ak[kk];
  return 0;
}

// SPECURV (src/mod_march.f90:2642-2689): the P2 Legendre moment of
// log10 E over the last fraction `ratio` of e[0..n-1], trapezoid rule;
// -0.05 when fewer than 10 points remain above 1e-22 or more than 20% of
// them were floored; clamped to [-10, 10]
static double specurv(const double *e, int n, double ratio) {
  int ie = n, is = ie - std::max(1, (int)(ie * ratio));
  while (ie > is && e[ie - 1] < 1.0e-22) ie--;
  const double rtrunc = 1.0 - (double)(ie - is) / (double)(n - is);
  const int nn = ie - is + 1;
  if (nn < 10 || rtrunc > 0.2) return -0.05;
  double curv = 0.0;
  const double dx = 2.0 / (nn - 1);
  for (int i = 1; i <= nn; i++) {
    const double xv = -1.0 + 2.0 * (i - 1) / (nn - 1);
    const double fac = (i == 1 || i == nn) ? 0.5 : 1.0;
    curv = curv + fac * std::log10(e[is + i - 2]) * dx * std::sqrt(5.0 / 8.0) * (3.0 * (xv * xv) - 1.0);
  }
  return std::fmin(10.0, std::fmax(-10.0, curv));
}

// SPESLOP (src/mod_march.f90:2729-2800): the P1 moment of log10 E against
// log10 k over the tail (k the 1-based index), floored at 1e-24 and at the
// tail's peak / 1e10, scaled to a slope; -3 with fewer than 3 points;
// clamped to [-15, 5]
static double speslop(const double *e, int n, double ratio) {
  int ie = n, is = ie - std::max(1, (int)(ie * ratio));
  const int k0 = is;
  double peak = e[is - 1];
  for (int i = is; i <= ie; i++) peak = std::fmax(peak, e[i - 1]);
  const double rel = peak / 1.0e10;
  while (ie > is && (e[ie - 1] < 1.0e-24 || e[ie - 1] < rel)) ie--;
  const int nn = ie - is + 1;
  if (nn < 3) return -3.0;
  const double xi0 = std::log10((double)k0), xi1 = std::log10((double)(k0 + nn - 1));
  double slope = 0.0;
  const double dx = 2.0 / (nn - 1);
  for (int i = 1; i <= nn; i++) {
    const double xi = std::log10((double)(k0 + i - 1));
    const double xv = -1.0 + 2.0 * (xi - xi0) / (xi1 - xi0);
    const double fac = (i == 1 || i == nn) ? 0.5 : 1.0;
    slope = slope + fac * std::log10(e[is + i - 2]) * dx * std::sqrt(3.0 / 2.0) * xv;
  }
  slope = slope * 2.0 / (xi1 - xi0);
  return std::fmin(5.0, std::fmax(-15.0, slope));
}

// PDAJUST (src/mod_march.f90:2841-2867): control = KP (cur - prev) + KD cur
// + OFFSET (with a threshold: forced to -(0.1 + 0.05 min(5, |cur - thr|))
// when cur and prev are both below it), clamped to [-0.4, 1] after SCALE;
// coeff = (1 + control) coeff, not below the baseline
static void pdajust(double &coeff, double cur, double prev, double base, const double *p, bool has_thr, double thr) {
  double control = p[0] * (cur - prev) + p[1] * cur + p[2];
  if (has_thr && cur < thr && prev < thr) control = -(0.1 + 0.05 * std::fmin(5.0, std::fabs(cur - thr)));
  control = std::fmin(1.0, std::fmax(-0.4, p[3] * control));
  coeff = (1.0 + control) * coeff;
  if (coeff < base) coeff = base;
}

// HYP3ADJ's two controllers (curvature: [2, 1, 0.05, 0.8], threshold -1;
// slope: [1.5, 0.5, 1.5, 1]) on one coefficient set c = (R, T, X) with its
// history h = (CURVMO, CURVKO, SLOPMO, SLOPKO); T and X take the larger of
// the two controllers' values, R the larger of the two on the 3:1 weighted
// mean of the m and k metrics, capped at 2 (src/mod_march.f90:3060-3110)
static void controllers(double *c, double *h, double base, double smc, double skc, double sms, double sks) {
  static const double cp[4] = {2.0, 1.0, 0.05, 0.8}, sp[4] = {1.5, 0.5, 1.5, 1.0};
  double tc = c[1], xc = c[2], ts = c[1], xs = c[2];
  pdajust(tc, smc, h[0], base, cp, true, -1.0);
  pdajust(xc, skc, h[1], base, cp, true, -1.0);
  pdajust(ts, sms, h[2], base, sp, false, 0.0);
  pdajust(xs, sks, h[3], base, sp, false, 0.0);
  c[1] = std::fmax(tc, ts);
  c[2] = std::fmax(xc, xs);
  const double rco = smc > skc ? (3.0 * h[0] + h[1]) / 4.0 : (h[0] + 3.0 * h[1]) / 4.0;
  const double rcn = smc > skc ? (3.0 * smc + skc) / 4.0 : (smc + 3.0 * skc) / 4.0;
  const double rso = sms > sks ? (3.0 * h[2] + h[3]) / 4.0 : (h[2] + 3.0 * h[3]) / 4.0;
  const double rsn = sms > sks ? (3.0 * sms + sks) / 4.0 : (sms + 3.0 * sks) / 4.0;
  double rc = c[0], rs = c[0];
  pdajust(rc, rcn, rco, base, cp, true, -1.0);
  pdajust(rs, rsn, rso, base, sp, false, 0.0);
  c[0] = std::fmin(std::fmax(rc, rs), 2.0);
  h[0] = smc;
  h[1] = skc;
  h[2] = sms;
  h[3] = sks;
}

int upload_hyperv(mlg_ctx *c) {
  Grid &G = c->g;
  hyp3set(G, c->hv_coef[0], c->hv_coef[1], c->hv_coef[2], G.hv_r, G.hv_t, G.hv_x);
  hyp3set(G, c->hv_coef[3], c->hv_coef[4], c->hv_coef[5], G.hd_r, G.hd_t, G.hd_x);
  if (!c->d.hv) return 0;  // no filter in this run: the tables are kept on the host only
  std::vector<double> hv;
  for (const std::vector<double> *v : {&G.hv_r, &G.hv_t, &G.hv_x, &G.hd_r, &G.hd_t, &G.hd_x})
    hv.insert(hv.end(), v->begin(), v->end());
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(c->d.hv, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice));
  return 0;
}

// HYP3ADJ (src/mod_march.f90:3000-3182) and the two HYP3SETs TIME_STEPPING
// makes after it (:282-287)
static int hyp3adj(mlg_ctx *c, const mlg_field *psi, const mlg_field *chi, const mlg_field *b) {
  const Grid &G = c->g;
  const int nt = G.ntchop, nxc = G.nxchop;
  std::vector<double> em(nt), ek(nxc);
  RET(mlg_chopset(c, 2));
  int rc = enemon(psi, chi, em.data(), ek.data());
  mlg_chopset(c, -2);
  RET(rc);
  {
    const double smc = specurv(em.data(), nt - 2, 1.0 / 4.0), skc = specurv(ek.data(), nxc, 1.0 / 4.0);
    const double sms = speslop(em.data(), nt - 2, 1.0 / 3.0), sks = speslop(ek.data(), nxc, 1.0 / 3.0);
    if (c->nupb < 0.0) {
      c->hist[0] = smc, c->hist[1] = skc, c->hist[2] = sms, c->hist[3] = sks;
      c->nupb = 0.8;
    }
    controllers(c->hv_coef, c->hist, c->nupb, smc, skc, sms, sks);
  }
  if (b) {
    // CALC_BOUSSI_PE (src/mod_boussinesq.f90:112-135) at the current truncation
    Temps T(c);
    mlg_field *w = T.get();
    RET(T.ready());
    RET(copy_field(w, b));
    RET(launch_set_ln(c, w->ln, nullptr, 0.0, 0.0));
    RET(mlg_rtran(w, 1));
    std::vector<double> pe;
    RET(halfk(c, w->e, nullptr, 1, pe));
    for (double &v : pe) v = v / 2.0 / (G.p.bv0 * G.p.bv0);
    std::vector<double> pm(nt), pk(nxc);
    sums(pe, nt, nxc, nt, pm.data(), pk.data());
    const double smc = specurv(pm.data(), nt - 2, 1.0 / 4.0), skc = specurv(pk.data(), nxc, 1.0 / 4.0);
    const double sms = speslop(pm.data(), nt - 2, 1.0 / 3.0), sks = speslop(pk.data(), nxc, 1.0 / 3.0);
    if (c->kappapb < 0.0) {
      c->phist[0] = smc, c->phist[1] = skc, c->phist[2] = sms, c->phist[3] = sks;
      c->kappapb = c->nupb;
    }
    controllers(c->hv_coef + 3, c->phist, c->kappapb, smc, skc, sms, sks);
  }
  return upload_hyperv(c);
}

// REMOVE (src/mod_march.f90:2503-2551): psi through DEL^2_H, chi through
// DEL^2 at NRCHOP+2, b directly, each with RINGCUT in PFF between RTRAN
// backward and forward (NS = 2 NR / 3, NS0 = min(NS + 4, NR))
static int remove_impl(mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = psi->ctx;
  const int ns = c->g.nr * 2 / 3, ns0 = std::min(ns + 4, c->g.nr);
  Temps T(c);
  mlg_field *wk = T.get();
  RET(T.ready());
  RET(mlg_delsqh(psi, wk));
  RET(mlg_rtran(wk, 1));
  RET(ringcut(wk, ns0, ns, 0));
  RET(mlg_rtran(wk, -1));
  RET(mlg_idelsqh(wk, psi));
  RET(mlg_chopset(c, 2));
  int rc = mlg_del2(chi, wk);
  if (!rc) rc = mlg_rtran(wk, 1);
  if (!rc) rc = ringcut(wk, ns0, ns, 1);
  mlg_chopset(c, -2);
  RET(rc);
  RET(mlg_rtran(wk, -1));
  double ln = 0.0;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(&ln, chi->ln, sizeof(double), hipMemcpyDeviceToHost));
  RET(mlg_idel2(wk, chi, 1, ln));
  if (b) {
    RET(mlg_rtran(b, 1));
    RET(ringcut(b, ns0, ns, 1));
    RET(mlg_rtran(b, -1));
  }
  return 0;
}

}  // namespace mlg

// ---------------------------------------------------------------- C ABI
extern "C" int mlg_ringcut(mlg_field *wk, int ns0, int ns, int divxm2) {
  if (!wk) return MLG_E_ARG;
  RET(ringcut(wk, ns0, ns, divxm2 ? 1 : 0));
  return mlg_synchronize(wk->ctx);
}

extern "C" int mlg_remove(mlg_field *psi, mlg_field *chi, mlg_field *b) {
  if (!psi || !chi || chi->ctx != psi->ctx || (b && b->ctx != psi->ctx)) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF || (b && b->space != MLG_FFF)) return MLG_E_SPACE;
  RET(remove_impl(psi, chi, b));
  return mlg_synchronize(psi->ctx);
}

extern "C" int mlg_sponge(mlg_field *psi, mlg_field *chi, mlg_field *b, double dt) {
  if (!psi || !chi || chi->ctx != psi->ctx || (b && b->ctx != psi->ctx)) return MLG_E_ARG;
  mlg_ctx *c = psi->ctx;
  if (c->rmv_sw != 2) return 0;
  if (psi->space != MLG_FFF) return 2;  // ERR_FLAGS%TRANSFORM
  // PSI%LN and CHI%LN are set aside and restored (mod_march.f90:2397-2398,2419-2420)
  double ln[2] = {0.0, 0.0};
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(&ln[0], psi->ln, sizeof(double), hipMemcpyDeviceToHost));
  MLG_CHECK_HIP(hipMemcpy(&ln[1], chi->ln, sizeof(double), hipMemcpyDeviceToHost));
  RET(launch_set_ln(c, psi->ln, nullptr, 0.0, 0.0));
  RET(launch_set_ln(c, chi->ln, nullptr, 0.0, 0.0));
  mlg_field *const f[3] = {psi, chi, b};
  RET(sponge_fields(c, f, b ? 3 : 2, dt));
  RET(launch_set_ln(c, psi->ln, nullptr, 0.0, ln[0]));
  RET(launch_set_ln(c, chi->ln, nullptr, 0.0, ln[1]));
  return mlg_synchronize(c);
}

extern "C" int mlg_enemon(const mlg_field *psi, const mlg_field *chi, double *em, double *ek) {
  if (!psi || !chi || !em || !ek || chi->ctx != psi->ctx) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return MLG_E_SPACE;
  RET(enemon(psi, chi, em, ek));
  return mlg_synchronize(psi->ctx);
}

extern "C" int mlg_hyp3adj(mlg_ctx *c, const mlg_field *psi, const mlg_field *chi, const mlg_field *b) {
  if (!c || !psi || !chi || psi->ctx != c || chi->ctx != c || (b && b->ctx != c)) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF || (b && b->space != MLG_FFF)) return MLG_E_SPACE;
  RET(hyp3adj(c, psi, chi, b));
  return mlg_synchronize(c);
}

extern "C" int mlg_hyperv_coeffs(const mlg_ctx *c, double out[6]) {
  if (!c || !out) return MLG_E_ARG;
  for (int i = 0; i < 6; i++) out[i] = c->hv_coef[i];
  return 0;
}

extern "C" int mlg_set_hyperv_coeffs(mlg_ctx *c, const double in[6]) {
  if (!c || !in) return MLG_E_ARG;
  for (int i = 0; i < 6; i++) c->hv_coef[i] = in[i];
  return upload_hyperv(c);
}

extern "C" int mlg_set_march_controls(mlg_solver *s, int rmv_sw, int rmv_int, int adj_sw, int adj_int) {
  if (!s || rmv_int < 0 || adj_int < 0) return MLG_E_ARG;
  s->rmv_sw = rmv_sw;
  s->rmv_int = rmv_int;
  s->adj_sw = adj_sw;
  s->adj_int = adj_int;
  return sponge_init(s->ctx, rmv_sw);
}

extern "C" int mlg_sponge_table(const mlg_ctx *c, double *tophat, int n) {
  if (!c || !tophat || n < c->g.ndimx) return MLG_E_ARG;
  for (int k = 0; k < c->g.ndimx; k++) tophat[k] = c->tophat.empty() ? 0.0 : c->tophat[k];
  return 0;
}
