// DIAGNOST's every-IT1 section (src/mod_march.f90:1797-1961) and CALCXY
// (:1967-2047): the domain integrals of vorticity, axial momentum, energy,
// helicity and angular momentum, the spread of psi and chi at r = infinity
// (CALCAT1), and the enstrophy-weighted vortex centre and radius.
//
// The fields are formed by the library's operators at NRCHOP+3; the
// integrals are the first FFF coefficient (INTEG, src/mod_scalar3.f90:
// 312-353) or Gauss sums over PFF columns (PRODCT :398-468, the per-(m, k)
// product launch of the energy diagnostics); the two physical-space
// products of CALCXY (ABSVEC2 :2050-2099, MULRCOSTH / MULRSINTH / MULR2
// :2102-2225) are pointwise kernels on PPP.
#include <cmath>
#include <vector>

#include "mlegs_internal.h"

using namespace mlg;

#define RET(x)                 \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

namespace mlg {

// ABSVEC2: a = ((u^2 + v^2) / r^2 + w^2) / (1-x)^2 on each theta sub-grid
// (the real and imaginary slots), r < NR, theta < NTH; a may be w
__global__ void __launch_bounds__(256) k_absvec2(KGrid g, const double2 *u, const double2 *v, const double2 *w, double2 *a,
                                                 const double *__restrict__ r2, const double *__restrict__ x) {
  const long n = (long)g.nr * g.nth * g.ndimx;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const int r = (int)(i % g.nr);
    const long rest = i / g.nr;
    const int j = (int)(rest % g.nth), z = (int)(rest / g.nth);
    const size_t at = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    const double2 uu = u[at], vv = v[at], ww = w[at];
    const double om = 1 - x[r], f = om * om;
    const double aaa = ((uu.x * uu.x + vv.x * vv.x) / r2[r] + ww.x * ww.x);
    const double bbb = ((uu.y * uu.y + vv.y * vv.y) / r2[r] + ww.y * ww.y);
    a[at] = make_double2(aaa / f, bbb / f);
  }
}

// a(r, theta, z) times fr(r, theta) (real slot) and fi(r, theta) (imaginary
// slot), r < NR, theta < NTH; fr, fi [NTH][NR]
__global__ void __launch_bounds__(256) k_mul_rtheta(KGrid g, double2 *a, const double *__restrict__ fr,
                                                    const double *__restrict__ fi) {
  const long n = (long)g.nr * g.nth * g.ndimx;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const int r = (int)(i % g.nr);
    const long rest = i / g.nr;
    const int j = (int)(rest % g.nth), z = (int)(rest / g.nth);
    const size_t at = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    const size_t t = (size_t)r + (size_t)g.nr * j;
    const double2 v = a[at];
    a[at] = make_double2(fr[t] * v.x, fi[t] * v.y);
  }
}

static int blocks_ppp(const KGrid &g) {
  const long n = (long)g.nr * g.nth * g.ndimx;
  return (int)std::min<long>((n + 255) / 256, 4096);
}

// INTEG: 4 pi ZLEN ELL^2 Re(f(0,0,0)) NORM(0, m=0) of an FFF field
static int integ(mlg_ctx *c, const mlg_field *f, double *out) {
  if (f->space != MLG_FFF) return 2;
  double2 e0;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(&e0, f->e, sizeof(double2), hipMemcpyDeviceToHost));
  const double pi = 3.14159265358979323846;
  *out = 4.0 * pi * c->g.zlen * c->g.ell2 * e0.x * c->g.norm[0];
  return 0;
}

// PRODCT: 2 pi ZLEN ELL^2 sum over r, m, k of w_r Re(a conj b), the modes
// m >= 1 twice (the per-(m, k) sums with m = 0 halved, added up)
static int prodct(mlg_ctx *c, const mlg_field *a, const mlg_field *b, double *out) {
  const Grid &G = c->g;
  const size_t na = (size_t)G.ntchop * G.nxchopdim;
  RET(launch_energy_terms(c, a->e, b->e, 0, c->d.red));
  std::vector<double> h(na);
  MLG_CHECK_HIP(hipMemcpyAsync(h.data(), c->d.red, sizeof(double) * na, hipMemcpyDeviceToHost, c->stream));
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double v : h) s = s + v;
  *out = s;
  return 0;
}

// CALCAT1 (src/mod_scalar3.f90:286-310): per k, sum over n < NRCHOP of
// f(n, m=0, k) AT1(n); returns |value at k = 0| and max - min of |value|
static int calcat1(mlg_ctx *c, const mlg_field *f, double *at0, double *spread) {
  const Grid &G = c->g;
  const int nn = G.nrchop, nxd = G.nxchopdim;
  std::vector<double2> col((size_t)nn * nxd);
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy2D(col.data(), sizeof(double2) * nn, f->e, sizeof(double2) * (size_t)G.nrchopdim * G.ntchopdim,
                            sizeof(double2) * nn, nxd, hipMemcpyDeviceToHost));
  double mx = -1.0, mn = 1e300;
  for (int k = 0; k < nxd; k++) {
    double re = 0.0, im = 0.0;
    for (int n = 0; n < nn; n++) {
      re = re + col[(size_t)nn * k + n].x * G.at1[n];
      im = im + col[(size_t)nn * k + n].y * G.at1[n];
    }
    const double a = std::hypot(re, im);
    if (k == 0) *at0 = a;
    mx = std::max(mx, a);
    mn = std::min(mn, a);
  }
  *spread = mx - mn;
  return 0;
}

static int diagnost_it1(const mlg_field *psi, const mlg_field *chi, double *out) {
  mlg_ctx *c = psi->ctx;
  const Grid &G = c->g;
  const double pi = 3.14159265358979323846;
  Temps T(c);
  mlg_field *rur = T.get(), *rup = T.get(), *uzh = T.get(), *ror = T.get(), *rop = T.get(), *ozh = T.get();
  RET(T.ready());
  // PC2VEL(..., 'H'): only UZ/(1-x)^2 = -DELSQH(chi) is read (INTEG); the
  // minus sign takes the elements, not LN
  RET(mlg_delsqh(chi, uzh));
  RET(launch_negate(c, uzh->e, c->n_fff));
  // PC2VOR(..., 'H'): ROR, ROP of PC2VOR and OZ/(1-x)^2 = -DELSQH(psi)
  RET(mlg_pc2vor(psi, chi, ror, rop, ozh));
  RET(mlg_delsqh(psi, ozh));
  RET(launch_negate(c, ozh->e, c->n_fff));
  RET(integ(c, ozh, &out[0]));  // VALVOR
  RET(integ(c, uzh, &out[1]));  // VALZMO
  // VALANG = -INTEG_P(XXDX(psi with LN = 0)).  INTEG_P (src/mod_march.f90:
  // 3649-3682) sums Re(F%E(:NR,1,1)) w / (1-x)^2 of its FFF argument: the
  // first NR stored coefficients in storage order (past the m = 0, k = 0
  // column when NRCHOPDIM < NR), as the reference reads them
  RET(copy_field(rur, psi));
  RET(launch_set_ln(c, rur->ln, nullptr, 0.0, 0.0));
  RET(mlg_xxdx(rur, rup));
  {
    std::vector<double2> e((size_t)G.nr);
    MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
    MLG_CHECK_HIP(hipMemcpy(e.data(), rup->e, sizeof(double2) * G.nr, hipMemcpyDeviceToHost));
    double dot = 0.0;
    for (int i = 0; i < G.nr; i++) {
      const double om = 1 - G.x[i];
      dot = dot + e[i].x / (om * om) * G.w[i];
    }
    out[4] = -(2.0 * pi * G.zlen * G.ell2 * dot);
  }
  // VALENE = -PRODCT(DELSQH(psi, 2 LN), psi) + PRODCT(DELSQH(chi), DEL2(chi)),
  // VALHEL = 2 PRODCT(DELSQH(psi), DEL2(chi)), all in PFF
  double ene1, ene2, hel1;
  RET(copy_field(rup, psi));
  RET(launch_set_ln(c, rup->ln, psi->ln, 2.0, 0.0));
  RET(mlg_delsqh(rup, rur));
  RET(mlg_rtran(rur, 1));
  RET(copy_field(rup, psi));
  RET(launch_set_ln(c, rup->ln, nullptr, 0.0, 0.0));
  RET(mlg_rtran(rup, 1));
  RET(prodct(c, rur, rup, &ene1));
  RET(mlg_del2(chi, rup));
  RET(mlg_rtran(rup, 1));
  RET(mlg_delsqh(chi, rur));
  RET(mlg_rtran(rur, 1));
  RET(prodct(c, rur, rup, &ene2));
  out[2] = -ene1 + ene2;
  RET(mlg_delsqh(psi, rur));
  RET(mlg_rtran(rur, 1));
  RET(prodct(c, rur, rup, &hel1));
  out[3] = 2 * hel1;
  // psi and chi at r = infinity
  RET(calcat1(c, psi, &out[5], &out[6]));
  RET(calcat1(c, chi, &out[7], &out[8]));
  // OZH (1-x)^2 = OZ, then CALCXY
  RET(mlg_mulxm(ozh, rur));
  RET(mlg_mulxm(rur, ozh));
  RET(mlg_tofp(ror));
  RET(mlg_tofp(rop));
  RET(mlg_tofp(ozh));
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_absvec2, dim3(blocks_ppp(g)), dim3(256), 0, c->stream, g, ror->e, rop->e, ozh->e, ozh->e, c->d.r,
                     c->d.x);
  MLG_CHECK_HIP(hipGetLastError());
  // the (r, theta) factors r cos(theta), r sin(theta), r^2 on both sub-grids
  // (THR = 2 pi j / NTH, THI = 2 pi (j + 1/2) / NTH; MINC = 1)
  const size_t nt = (size_t)G.nr * G.nth;
  std::vector<double> f(6 * nt);
  for (int j = 0; j < G.nth; j++) {
    const double thr = 2 * pi / G.p.minc / G.nth * j, thi = 2 * pi / G.p.minc / G.nth * (j + 0.5);
    for (int r = 0; r < G.nr; r++) {
      const size_t t = (size_t)r + (size_t)G.nr * j;
      f[t] = G.r[r] * std::cos(thr);
      f[nt + t] = G.r[r] * std::cos(thi);
      f[2 * nt + t] = G.r[r] * std::sin(thr);
      f[3 * nt + t] = G.r[r] * std::sin(thi);
      f[4 * nt + t] = G.r[r] * G.r[r];
      f[5 * nt + t] = G.r[r] * G.r[r];
    }
  }
  double *df = nullptr;
  MLG_CHECK_HIP(hipMalloc((void **)&df, sizeof(double) * f.size()));
  int rc = hipMemcpy(df, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice) == hipSuccess ? 0 : MLG_E_HIP;
  double valens = 0.0, mom[3] = {0.0, 0.0, 0.0};
  if (!rc) rc = copy_field(ror, ozh);
  if (!rc) rc = mlg_toff(ror);
  if (!rc) rc = integ(c, ror, &valens);
  for (int q = 0; q < 3 && !rc; q++) {
    rc = copy_field(ror, ozh);
    if (!rc) {
      hipLaunchKernelGGL(k_mul_rtheta, dim3(blocks_ppp(g)), dim3(256), 0, c->stream, g, ror->e, df + 2 * q * nt,
                         df + (2 * q + 1) * nt);
      if (hipGetLastError() != hipSuccess) rc = MLG_E_HIP;
    }
    if (!rc) rc = mlg_toff(ror);
    if (!rc) rc = integ(c, ror, &mom[q]);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(df);
  RET(rc);
  out[9] = mom[0] / valens;             // <X>
  out[10] = mom[1] / valens;            // <Y>
  out[11] = std::sqrt(mom[2] / valens); // <R>
  return 0;
}

}  // namespace mlg

// out: VALVOR, VALZMO, VALENE, VALHEL, VALANG, |PSIAT1(k=0)|, spread of
// |PSIAT1|, |CHIAT1(k=0)|, spread of |CHIAT1|, <X>, <Y>, <R>
extern "C" int mlg_diagnost_it1(const mlg_field *psi, const mlg_field *chi, double out[12]) {
  if (!psi || !chi || !out || chi->ctx != psi->ctx) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return MLG_E_SPACE;
  mlg_ctx *c = psi->ctx;
  RET(mlg_chopset(c, 3));
  const int rc = diagnost_it1(psi, chi, out);
  mlg_chopset(c, -3);
  RET(rc);
  return mlg_synchronize(c);
}
