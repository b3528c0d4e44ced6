// Discrete adjoint of the AB2-CN Boussinesq march: the exact transpose of
// what mlg_initialize (RICH_BOUSSI_FE_BE, src/mod_march.f90:535-606) and
// mlg_time_step (STEP_BOUSSI_AB_CN, src/mod_march.f90:609-661, with
// CALC_BOUSSI_FORCE, src/mod_boussinesq.f90:317-410, and CALC_BOUSSI_VISC_CN
// / _BE, src/mod_march.f90:1200-1265) compute, for the inner product
// <a, b> = sum over stored coefficients of Re(conj(a) b) + a%LN b%LN (psi
// and chi carry a log term; b's is not a degree of freedom).
//
// The step is quadratic in the state only through the products of
// CALC_BOUSSI_FORCE, so its transpose at a state needs that state's
// physical velocity, vorticity and buoyancy gradient, which are recomputed
// here from the state (the caller keeps states, not products: the
// trajectory is checkpointed).  Every other stage is linear and is
// transposed operator by operator:
//   RTRAN, VERFFT, HORFFT      the opposite transform between diagonal scalings
//   PROJECT, VEL2PLN, CALCAT1  k_adj_project, k_adj_vel2pln, k_adj_fixinf
//   PC2VEL, PC2VOR, GRAD       k_adj_columns
//   IDEL2LN, IHELM             substitutions with the transposed cached factors
//   DEL2                       k_adj_del2t
// The reference has no adjoint; nothing here replaces a reference routine.
#include <cmath>

#include "mlegs_internal.h"

using namespace mlg;

#define RET(x)                 \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

namespace {

// CHOPSET(iof) for the lifetime of the object
struct Chop {
  mlg_ctx *c;
  int iof;
  Chop(mlg_ctx *cc, int i) : c(cc), iof(i) { mlg_chopset(c, iof); }
  ~Chop() { mlg_chopset(c, -iof); }
};

void swap_e(mlg_field *a, mlg_field *b) { std::swap(a->e, b->e); }

// The adjoint supports what the AB2-CN march of these tests runs: Adams-
// Bashforth with Crank-Nicolson (or no) viscosity, no hyperviscosity, FP64
// Legendre products.
int check_supported(const mlg_ctx *c) {
  const mlg_params &p = c->g.p;
  if (!p.adams || p.hyper_cn || (p.visc_sw != 0 && p.visc_sw != 1)) return MLG_E_ARG;
  if (c->leg_split) return MLG_E_ARG;
  return 0;
}

// Fields of one batch swap storage with scratch after an out-of-place
// transform (only the element buffers move; spaces are set by the caller).
void swap_n(int n, mlg_field *const *a, mlg_field *const *t) {
  for (int i = 0; i < n; i++) swap_e(a[i], t[i]);
}

// transpose of TOFP (RTRAN, VERFFT, HORFFT backward) for n PPP adjoints:
// HORFFT bwd^T = 2 NTH D2 HORFFT fwd (D2 = 1 at m = 0, 2 above), VERFFT
// bwd^T = NX VERFFT fwd, RTRAN bwd^T = the Gauss sums without weights.
// nrchop[i]: the truncation of field i's RTRAN (its TOFP's); t: n scratch.
int tofp_adj_n(mlg_ctx *c, int n, mlg_field *const *f, mlg_field *const *t, const int *nrchop) {
  const Grid &G = c->g;
  double2 *e[MAX_HOR_JOBS];
  for (int i = 0; i < n; i++) e[i] = f[i]->e;
  const double s = 2.0 * G.nth * G.nx;
  RET(launch_horfft_fwd_n(c, n, e, s, 2.0 * s));
  for (int i0 = 0; i0 < n; i0 += MAX_TRANSFORM_JOBS) {
    const int m = n - i0 < MAX_TRANSFORM_JOBS ? n - i0 : MAX_TRANSFORM_JOBS;
    const double2 *in[MAX_TRANSFORM_JOBS];
    double2 *out[MAX_TRANSFORM_JOBS];
    for (int i = 0; i < m; i++) in[i] = f[i0 + i]->e, out[i] = t[i0 + i]->e;
    RET(launch_verfft_fwd_n(c, m, in, out));
  }
  swap_n(n, f, t);
  // RTRAN forward without weights, fields of one truncation per launch
  for (int i0 = 0; i0 < n;) {
    int m = 1;
    while (i0 + m < n && m < MAX_TRANSFORM_JOBS && nrchop[i0 + m] == nrchop[i0]) m++;
    const double2 *in[MAX_TRANSFORM_JOBS];
    double2 *out[MAX_TRANSFORM_JOBS];
    for (int i = 0; i < m; i++) in[i] = f[i0 + i]->e, out[i] = t[i0 + i]->e;
    const int d = nrchop[i0] - c->g.nrchop;
    if (d) RET(mlg_chopset(c, d));
    const int rc = launch_rtran_fwd_n(c, m, in, out, c->d.ones);
    if (d) mlg_chopset(c, -d);
    RET(rc);
    i0 += m;
  }
  for (int i = 0; i < n; i++) f[i]->space = MLG_FFF;
  swap_n(n, f, t);
  return 0;
}

// transpose of TOFF (HORFFT, VERFFT, RTRAN forward) for n FFF adjoints into
// the PPP fields out: RTRAN fwd^T = w RTRAN bwd (w folded into the HORFFT
// load), VERFFT fwd^T = VERFFT bwd / NX, HORFFT fwd^T = HORFFT bwd D(1/2) /
// (2 NTH).  t: n scratch; zln: an LN of value 0.
int toff_adj_n(mlg_ctx *c, int n, const mlg_field *const *in, mlg_field *const *out, mlg_field *const *t, const double *zln) {
  const Grid &G = c->g;
  const double2 *ie[MAX_TRANSFORM_JOBS];
  const double *lz[MAX_TRANSFORM_JOBS];
  double2 *te[MAX_TRANSFORM_JOBS], *oe[MAX_TRANSFORM_JOBS];
  if (n > MAX_TRANSFORM_JOBS) return MLG_E_ARG;
  for (int i = 0; i < n; i++) ie[i] = in[i]->e, lz[i] = zln, te[i] = t[i]->e, oe[i] = out[i]->e;
  RET(launch_rtran_bwd_n(c, n, ie, lz, te));
  const double2 *tc[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) tc[i] = te[i];
  RET(launch_verfft_bwd_n(c, n, tc, oe));
  const double s = 1.0 / (2.0 * G.nth * G.nx);
  RET(launch_horfft_bwd_n(c, n, oe, s, 0.5 * s, c->d.w));
  for (int i = 0; i < n; i++) out[i]->space = MLG_PPP;
  return 0;
}

// transpose of PROJECT from PPP with VEL2PLN (psi%LN set, as
// CALC_BOUSSI_FORCE calls it): psi, chi2 adjoints (FFF) into three PPP
// adjoints; v: three PFF scratch fields
int project_adj(mlg_ctx *c, const mlg_field *psib, const mlg_field *chib, mlg_field *const *out, mlg_field *const *v) {
  RET(launch_adj_project(c, psib->e, chib->e, v[0]->e, v[1]->e, v[2]->e));
  RET(launch_adj_vel2pln(c, v[1]->e));
  const double2 *ve[3] = {v[0]->e, v[1]->e, v[2]->e};
  double2 *oe[3] = {out[0]->e, out[1]->e, out[2]->e};
  RET(launch_verfft_bwd_n(c, 3, ve, oe));
  const double s = 1.0 / (2.0 * c->g.nth * c->g.nx);
  RET(launch_horfft_bwd_n(c, 3, oe, s, 0.5 * s, nullptr));
  for (int i = 0; i < 3; i++) out[i]->space = MLG_PPP;
  return 0;
}

// x <- IHELM^T(x) with LN, at the current truncation
int ihelm_adj(mlg_field *x, double alp) {
  BandTJob j{x->e, x->e, x->ln, x->ln, alp, nullptr};
  return launch_band_solve_t_n(x->ctx, 2, 1, &j);
}

// x <- VISC1^T(x): VISC1 is A = IHELM(-ALP A, ALP), ALP = 1/(NU DT)
int visc1_adj(mlg_field *x, double nu, double dt) {
  const double alp = 1 / (nu * dt);
  RET(ihelm_adj(x, alp));
  RET(launch_adj_lincomb(x->ctx, x->e, -alp, nullptr, 0.0, nullptr, 0.0, x->ctx->n_fff));
  return launch_adj_ln_lincomb(x->ctx, x->ln, -alp, nullptr, 0.0, nullptr, 0.0);
}

// Transpose of CALC_BOUSSI_FORCE at the state (psi, chi, b): the adjoints
// of its results (psin, chin, bn: FFF) are carried back and added to
// psib, chib, bb.
int force_adj(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, const mlg_field *psinb, const mlg_field *chinb,
              const mlg_field *bnb, mlg_field *psib, mlg_field *chib, mlg_field *bb) {
  mlg_ctx *c = psi->ctx;
  Temps T(c);
  mlg_field *u[10], *t[10];
  for (auto &f : u) f = T.get();  // u[0..8]: u, om, grad b; u[9]: the adjoint of bn's and W's physical field
  for (auto &f : t) f = T.buffer();
  mlg_field *w2 = T.buffer(), *z = T.get();
  RET(T.ready());
  const int nrchop = c->g.nrchop;
  // the products' inputs at the state, in physical space (CHOPSET(3) as in
  // CALC_BOUSSI_FORCE): PC2VEL + PC2VOR, GRAD, then TOFP of all nine
  {
    Chop ch(c, 3);
    VelVorOut vo;
    for (int i = 0; i < 6; i++) vo.e[i] = u[i]->e, vo.ln[i] = u[i]->ln;
    RET(launch_pc2velvor(c, psi->e, psi->ln, chi->e, chi->ln, vo));
    RET(launch_grad(c, b->e, b->ln, u[6]->e, u[6]->ln, u[7]->e, u[7]->ln, u[8]->e, u[8]->ln));
    const double2 *in[9];
    const double *ln[9];
    double2 *out[9];
    for (int i = 0; i < 9; i++) in[i] = u[i]->e, ln[i] = u[i]->ln, out[i] = t[i]->e;
    RET(launch_rtran_bwd_n(c, 9, in, ln, out));
    swap_n(9, u, t);
    for (int i = 0; i < 9; i++) in[i] = u[i]->e, out[i] = t[i]->e;
    RET(launch_verfft_bwd_n(c, 9, in, out));
    swap_n(9, u, t);
    double2 *ue9[9];
    for (int i = 0; i < 9; i++) ue9[i] = u[i]->e;
    RET(launch_horfft_bwd_n(c, 9, ue9, 1.0, 1.0, nullptr));
    for (int i = 0; i < 9; i++) u[i]->space = MLG_PPP;
  }
  // bn = TOFF(bn_phys)
  RET(toff_adj_n(c, 1, &bnb, &u[9], &t[0], z->ln));
  // chin = IDEL2LN(w2), then CHOPDO and the fix at infinity
  BandTJob j{chinb->e, w2->e, chinb->ln, nullptr, 0.0, nullptr};
  RET(launch_band_solve_t_n(c, 1, 1, &j));
  // (psin, w2) = PROJECT(N_r, N_theta, N_z) from PPP, psi's fix at infinity last
  mlg_field *ob[3] = {t[3], t[4], t[5]};
  RET(project_adj(c, psinb, w2, ob, &t[0]));
  // the pointwise terms
  double2 *ue[9];
  for (int i = 0; i < 9; i++) ue[i] = u[i]->e;
  const double2 *oe[3] = {ob[0]->e, ob[1]->e, ob[2]->e};
  RET(launch_adj_products(c, ue, oe, u[9]->e));
  // TOFP^T of the nine (at NRCHOP+3) and of W (at NRCHOP)
  int chops[10];
  for (int i = 0; i < 9; i++) chops[i] = nrchop + 3;
  chops[9] = nrchop;
  RET(tofp_adj_n(c, 10, u, t, chops));
  RET(launch_adj_lincomb(c, bb->e, 1.0, u[9]->e, 1.0, nullptr, 0.0, c->n_fff));
  // PC2VEL, PC2VOR, GRAD at NRCHOP+3, into psi, chi, b rows of NRCHOP
  {
    Chop ch(c, 3);
    const double2 *ie[9];
    for (int i = 0; i < 9; i++) ie[i] = u[i]->e;
    RET(launch_adj_columns(c, nrchop, ie, psib->e, psib->ln, chib->e, chib->ln, bb->e));
  }
  return 0;
}

int fff_ok(const mlg_field *f) { return f && f->space == MLG_FFF; }

}  // namespace

extern "C" int mlg_force_adjoint(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, const mlg_field *psin_bar,
                                 const mlg_field *chin_bar, const mlg_field *bn_bar, mlg_field *psi_bar, mlg_field *chi_bar,
                                 mlg_field *b_bar) {
  if (!fff_ok(psi) || !fff_ok(chi) || !fff_ok(b) || !fff_ok(psin_bar) || !fff_ok(chin_bar) || !fff_ok(bn_bar) ||
      !fff_ok(psi_bar) || !fff_ok(chi_bar) || !fff_ok(b_bar))
    return MLG_E_ARG;
  mlg_ctx *c = psi->ctx;
  if (c->leg_split) return MLG_E_ARG;
  RET(force_adj(psi, chi, b, psin_bar, chin_bar, bn_bar, psi_bar, chi_bar, b_bar));
  return mlg_synchronize(c);
}

// Transpose of one STEP_BOUSSI_AB_CN from the state (psi, chi, b) (the
// history N_old enters linearly, so only the state is needed).  Forward:
//   N = FORCE(u); u* = u + dt (1.5 N - 0.5 N_old); N_old' = N;
//   u' = IHELM(-DEL2(u) - ALP u*, ALP), ALP = 2/(NU dt)   (CN, per field)
// In: ub = adjoint of u', hb = adjoint of N_old'.  Out: ub = adjoint of u,
// hb = adjoint of N_old.
extern "C" int mlg_step_adjoint(mlg_solver *s, const mlg_field *psi, const mlg_field *chi, const mlg_field *b,
                                mlg_field *psi_bar, mlg_field *chi_bar, mlg_field *b_bar, mlg_field *hpsi_bar,
                                mlg_field *hchi_bar, mlg_field *hb_bar) {
  if (!s || !fff_ok(psi) || !fff_ok(chi) || !fff_ok(b) || !fff_ok(psi_bar) || !fff_ok(chi_bar) || !fff_ok(b_bar) ||
      !fff_ok(hpsi_bar) || !fff_ok(hchi_bar) || !fff_ok(hb_bar))
    return MLG_E_ARG;
  mlg_ctx *c = s->ctx;
  RET(check_supported(c));
  const mlg_params &P = c->g.p;
  const double dt = P.dt;
  mlg_field *ub[3] = {psi_bar, chi_bar, b_bar}, *hb[3] = {hpsi_bar, hchi_bar, hb_bar};
  const double nu[3] = {P.nu, P.nu, P.kappa};
  const bool cn[3] = {P.visc_sw == 1, P.visc_sw == 1, P.visc_sw == 1 && P.kappa != 0.0};
  Temps T(c);
  mlg_field *wk[3] = {T.get(), T.get(), T.get()}, *nb[3] = {T.get(), T.get(), T.get()};
  RET(T.ready());
  // wk = IHELM^T(ubar) for the fields with CN, one launch
  BandTJob jobs[3];
  int nj = 0;
  double alp[3] = {0.0, 0.0, 0.0};
  for (int f = 0; f < 3; f++) {
    if (!cn[f]) continue;
    if (nu[f] == 0.0) return 5;
    alp[f] = 2 / nu[f] / dt;
    jobs[nj++] = BandTJob{ub[f]->e, wk[f]->e, ub[f]->ln, wk[f]->ln, alp[f], nullptr};
  }
  if (nj) RET(launch_band_solve_t_n(c, 2, nj, jobs));
  // ubar = u*bar - DEL2^T(wk), Nbar = CHOPDO(hbar) + 1.5 dt u*bar, hbar = -0.5 dt u*bar, one launch
  AdjCnField fl[3];
  for (int f = 0; f < 3; f++) {
    const bool ln = f < 2;
    fl[f] = AdjCnField{wk[f]->e, wk[f]->ln, ub[f]->e, hb[f]->e, nb[f]->e, ln ? ub[f]->ln : nullptr,
                       ln ? hb[f]->ln : nullptr, ln ? nb[f]->ln : nullptr, alp[f], cn[f] ? 1 : 0};
  }
  {
    const int nrchop = c->g.nrchop;
    Chop ch(c, 2);
    RET(launch_adj_cn(c, nrchop, 3, fl, dt));
  }
  RET(force_adj(psi, chi, b, nb[0], nb[1], nb[2], psi_bar, chi_bar, b_bar));
  double *lz[2] = {b_bar->ln, hb_bar->ln};
  RET(launch_zero_lns(c, lz, 2));
  return mlg_synchronize(c);
}

// Transpose of RICH_BOUSSI_FE_BE from the initial state u0 (psi, chi, b).
// Forward:  N0 = FORCE(u0);  a = BE(u0 + dt/2 N0, dt/2);  u2 = BE(a + dt/2 FORCE(a), dt/2)
//           u1 = BE(u0 + dt N0, dt);  u' = 2 u2 - u1;  N_old' = N0
// (BE = VISC_BE: A = IHELM(-ALP A, ALP), ALP = 1/(NU dt_sub)).  In: ub, hb
// the adjoints of u' and N_old'; out: ub the adjoint of u0, hb zero.
extern "C" int mlg_start_adjoint(mlg_solver *s, const mlg_field *psi, const mlg_field *chi, const mlg_field *b,
                                 mlg_field *psi_bar, mlg_field *chi_bar, mlg_field *b_bar, mlg_field *hpsi_bar,
                                 mlg_field *hchi_bar, mlg_field *hb_bar) {
  if (!s || !fff_ok(psi) || !fff_ok(chi) || !fff_ok(b) || !fff_ok(psi_bar) || !fff_ok(chi_bar) || !fff_ok(b_bar) ||
      !fff_ok(hpsi_bar) || !fff_ok(hchi_bar) || !fff_ok(hb_bar))
    return MLG_E_ARG;
  mlg_ctx *c = s->ctx;
  RET(check_supported(c));
  const mlg_params &P = c->g.p;
  const double dt = P.dt, hdt = 0.5 * dt;
  const long n = c->n_fff;
  mlg_field *ub[3] = {psi_bar, chi_bar, b_bar}, *hb[3] = {hpsi_bar, hchi_bar, hb_bar};
  const double nu[3] = {P.nu, P.nu, P.kappa};
  const bool be[3] = {P.visc_sw == 1, P.visc_sw == 1, P.visc_sw == 1 && P.kappa != 0.0};
  for (int f = 0; f < 3; f++)
    if (be[f] && nu[f] == 0.0) return 5;
  Temps T(c);
  mlg_field *a[3] = {T.get(), T.get(), T.get()}, *n0[3] = {T.get(), T.get(), T.get()};
  mlg_field *u2b[3] = {T.get(), T.get(), T.get()}, *u1b[3] = {T.get(), T.get(), T.get()};
  mlg_field *n0b[3] = {T.get(), T.get(), T.get()};
  RET(T.ready());
  // forward: N0 and the half-step state a
  RET(mlg_force(psi, chi, b, n0[0], n0[1], n0[2]));
  const mlg_field *u0[3] = {psi, chi, b};
  for (int f = 0; f < 3; f++) {
    RET(copy_field(a[f], u0[f]));
    RET(launch_adj_lincomb(c, a[f]->e, 1.0, n0[f]->e, hdt, nullptr, 0.0, n));
    if (f < 2) RET(launch_adj_ln_lincomb(c, a[f]->ln, 1.0, n0[f]->ln, hdt, nullptr, 0.0));
  }
  if (P.visc_sw == 1) {
    Temps W(c);
    mlg_field *wk = W.get();
    RET(W.ready());
    for (int f = 0; f < 3; f++) {
      if (!be[f]) continue;
      const double alp = 1 / (nu[f] * hdt);
      RET(copy_field(wk, a[f]));
      RET(launch_adj_lincomb(c, wk->e, -alp, nullptr, 0.0, nullptr, 0.0, n));
      RET(launch_adj_ln_lincomb(c, wk->ln, -alp, nullptr, 0.0, nullptr, 0.0));
      RET(mlg_ihelm(wk, a[f], alp));
    }
  }
  RET(launch_set_ln(c, a[2]->ln, nullptr, 0.0, 0.0));
  RET(launch_set_ln(c, b_bar->ln, nullptr, 0.0, 0.0));
  // adjoint: u2bar = 2 ubar, u1bar = -ubar
  for (int f = 0; f < 3; f++) {
    RET(copy_field(u2b[f], ub[f]));
    RET(launch_adj_lincomb(c, u2b[f]->e, 2.0, nullptr, 0.0, nullptr, 0.0, n));
    RET(launch_adj_ln_lincomb(c, u2b[f]->ln, 2.0, nullptr, 0.0, nullptr, 0.0));
    RET(copy_field(u1b[f], ub[f]));
    RET(launch_adj_lincomb(c, u1b[f]->e, -1.0, nullptr, 0.0, nullptr, 0.0, n));
    RET(launch_adj_ln_lincomb(c, u1b[f]->ln, -1.0, nullptr, 0.0, nullptr, 0.0));
    if (be[f]) {
      RET(visc1_adj(u1b[f], nu[f], dt));
      RET(visc1_adj(u2b[f], nu[f], hdt));
    }
    // u1 = u0 + dt N0 (before BE): u0bar = u1bar, N0bar = dt u1bar + hbar
    RET(copy_field(ub[f], u1b[f]));
    RET(copy_field(n0b[f], hb[f]));
    RET(launch_adj_lincomb(c, n0b[f]->e, 1.0, u1b[f]->e, dt, nullptr, 0.0, n));
    RET(launch_adj_ln_lincomb(c, n0b[f]->ln, 1.0, u1b[f]->ln, dt, nullptr, 0.0));
  }
  // u2 = a + dt/2 FORCE(a) (before BE): abar = u2bar + FORCE'(a)^T (dt/2 u2bar)
  {
    Temps W(c);
    mlg_field *fb[3] = {W.get(), W.get(), W.get()};
    RET(W.ready());
    for (int f = 0; f < 3; f++) {
      RET(copy_field(fb[f], u2b[f]));
      RET(launch_adj_lincomb(c, fb[f]->e, hdt, nullptr, 0.0, nullptr, 0.0, n));
      RET(launch_adj_ln_lincomb(c, fb[f]->ln, f < 2 ? hdt : 0.0, nullptr, 0.0, nullptr, 0.0));
    }
    RET(force_adj(a[0], a[1], a[2], fb[0], fb[1], fb[2], u2b[0], u2b[1], u2b[2]));
  }
  // a = BE(u0 + dt/2 N0, dt/2)
  for (int f = 0; f < 3; f++) {
    if (be[f]) RET(visc1_adj(u2b[f], nu[f], hdt));
    RET(launch_adj_lincomb(c, ub[f]->e, 1.0, u2b[f]->e, 1.0, nullptr, 0.0, n));
    RET(launch_adj_ln_lincomb(c, ub[f]->ln, 1.0, u2b[f]->ln, 1.0, nullptr, 0.0));
    RET(launch_adj_lincomb(c, n0b[f]->e, 1.0, u2b[f]->e, hdt, nullptr, 0.0, n));
    RET(launch_adj_ln_lincomb(c, n0b[f]->ln, 1.0, u2b[f]->ln, hdt, nullptr, 0.0));
  }
  RET(launch_set_ln(c, n0b[2]->ln, nullptr, 0.0, 0.0));
  RET(launch_set_ln(c, b_bar->ln, nullptr, 0.0, 0.0));
  RET(force_adj(psi, chi, b, n0b[0], n0b[1], n0b[2], psi_bar, chi_bar, b_bar));
  RET(launch_set_ln(c, b_bar->ln, nullptr, 0.0, 0.0));
  for (int f = 0; f < 3; f++) {
    RET(launch_adj_lincomb(c, hb[f]->e, 0.0, nullptr, 0.0, nullptr, 0.0, n));
    RET(launch_set_ln(c, hb[f]->ln, nullptr, 0.0, 0.0));
  }
  return mlg_synchronize(c);
}

// Gradient of J = sum over m >= mmin (all k) of CALC_BOUSSI_ENERGY's KE(m,k)
// (src/mod_boussinesq.f90:112-232) with respect to psi and chi, written to
// psi_bar and chi_bar.  With R = RTRAN backward at NRCHOP+3, W the Gauss
// weights times 4 pi ZLEN L^2 on the modes m >= mmin (mmin >= 1; m = 0 would
// bring LN and the halved weight in), A = R^T W R, Dh = DELSQH, L = DEL2:
//   J = -1/2 Re(psi^H A Dh psi) + 1/2 Re(chi^H L^T A Dh chi)
//   dJ/dpsi = -1/2 (A Dh + Dh A) psi,   dJ/dchi = 1/2 (L^T A Dh + Dh A L) chi.
extern "C" int mlg_ke_gradient(const mlg_field *psi, const mlg_field *chi, int mmin, mlg_field *psi_bar, mlg_field *chi_bar) {
  if (!fff_ok(psi) || !fff_ok(chi) || !psi_bar || !chi_bar || mmin != 1) return MLG_E_ARG;
  mlg_ctx *c = psi->ctx;
  if (c->leg_split) return MLG_E_ARG;
  Temps T(c);
  mlg_field *y[4] = {T.get(), T.get(), T.get(), T.get()};
  mlg_field *t = T.buffer(), *z = T.get();
  RET(T.ready());
  const double pi = 3.14159265358979323846;
  const double fac = 4.0 * pi * c->g.zlen * c->g.ell2;
  const long n = c->n_fff;
  psi_bar->space = chi_bar->space = MLG_FFF;
  {
    Chop ch(c, 3);
    // y0 = Dh psi, y1 = psi, y2 = Dh chi, y3 = L chi (no log terms: m = 0 is left out)
    RET(launch_delsqh(c, psi->e, z->ln, y[0]->e, y[0]->ln));
    RET(launch_copy(c, y[1]->e, psi->e, n));
    RET(launch_delsqh(c, chi->e, z->ln, y[2]->e, y[2]->ln));
    RET(launch_del2(c, chi->e, z->ln, y[3]->e, y[3]->ln));
    // y <- A y = R^T W R y
    for (auto *f : y) {
      RET(launch_rtran_bwd(c, f->e, z->ln, t->e));
      RET(launch_adj_scale_pff(c, t->e, 2, fac, 1));
      RET(launch_rtran_fwd(c, t->e, f->e, nullptr));
    }
    // psi_bar = -1/2 (A Dh psi + Dh A psi)
    RET(launch_delsqh(c, y[1]->e, z->ln, psi_bar->e, psi_bar->ln));
    RET(launch_adj_lincomb(c, psi_bar->e, -0.5, y[0]->e, -0.5, nullptr, 0.0, n));
    // chi_bar = 1/2 (L^T A Dh chi + Dh A L chi)
    RET(launch_delsqh(c, y[3]->e, z->ln, chi_bar->e, chi_bar->ln));
    RET(launch_adj_lincomb(c, chi_bar->e, 0.5, nullptr, 0.0, nullptr, 0.0, n));
    RET(launch_adj_del2t(c, c->g.nrchop, y[2]->e, 0.5, chi_bar->e, nullptr));
  }
  RET(launch_chopdo(c, psi_bar->e, MLG_FFF));
  RET(launch_chopdo(c, chi_bar->e, MLG_FFF));
  RET(launch_set_ln(c, psi_bar->ln, nullptr, 0.0, 0.0));
  RET(launch_set_ln(c, chi_bar->ln, nullptr, 0.0, 0.0));
  return mlg_synchronize(c);
}

// Transposes of single operators, for callers that compose their own
// adjoints (and for testing each against its forward routine):
//   0 TOFP^T    f[0]: PPP adjoint in, FFF adjoint out
//   1 TOFF^T    f[0]: FFF adjoint in, PPP adjoint out
//   2 PROJECT^T from PPP with VEL2PLN and psi%LN = 0 (as CALC_BOUSSI_FORCE
//               calls it): f[0], f[1] psi, chi adjoints in; f[2..4] the
//               three PPP adjoints out
//   3 IDEL2LN^T f[0] in (with LN), f[1] out
//   4 IHELM^T   f[0] in (with LN), f[1] out (with LN); alp
//   5 DEL2^T    f[0] in, f[1] out (with LN)
//   6 (PC2VEL, PC2VOR, GRAD)^T at the current truncation: f[0..8] the
//               adjoints of ru_r ru_t u_z ro_r ro_t o_z gb_r gb_t gb_z in;
//               added to f[9..11] (psi, chi, b adjoints)
//   7 CALC_BOUSSI_FORCE's pointwise terms^T: f[0..8] the forward u, om, gb
//               (PPP) in, their adjoints out; f[9..11] the adjoints of the
//               three vector results; f[12] the adjoint of N_b in, of B out
extern "C" int mlg_adjoint_op(int op, mlg_field *const *f, int nf, double alp) {
  static const int need[8] = {1, 1, 5, 2, 2, 2, 12, 13};
  if (op < 0 || op > 7 || !f || nf < need[op]) return MLG_E_ARG;
  for (int i = 0; i < need[op]; i++)
    if (!f[i]) return MLG_E_ARG;
  mlg_ctx *c = f[0]->ctx;
  if (c->leg_split) return MLG_E_ARG;
  Temps T(c);
  mlg_field *t = T.buffer(), *z = T.get();
  mlg_field *v[3] = {T.buffer(), T.buffer(), T.buffer()};
  RET(T.ready());
  switch (op) {
    case 0: {
      if (f[0]->space != MLG_PPP) return 2;
      const int ch = c->g.nrchop;
      RET(tofp_adj_n(c, 1, &f[0], &t, &ch));
      break;
    }
    case 1: {
      if (f[0]->space != MLG_FFF) return 2;
      RET(copy_field(v[0], f[0]));
      const mlg_field *in = v[0];
      RET(toff_adj_n(c, 1, &in, &f[0], &t, z->ln));
      break;
    }
    case 2:
      RET(project_adj(c, f[0], f[1], &f[2], v));
      for (int i = 0; i < 3; i++) RET(launch_set_ln(c, f[2 + i]->ln, nullptr, 0.0, 0.0));
      break;
    case 3: {
      BandTJob j{f[0]->e, f[1]->e, f[0]->ln, nullptr, 0.0, nullptr};
      RET(launch_band_solve_t_n(c, 1, 1, &j));
      f[1]->space = MLG_FFF;
      RET(launch_set_ln(c, f[1]->ln, nullptr, 0.0, 0.0));
      break;
    }
    case 4:
      if (alp == 0.0) return 3;
      RET(copy_field(f[1], f[0]));
      RET(ihelm_adj(f[1], alp));
      break;
    case 5:
      RET(launch_zero(c, f[1]->e, c->n_fff));
      RET(launch_set_ln(c, f[1]->ln, nullptr, 0.0, 0.0));
      f[1]->space = MLG_FFF;
      RET(launch_adj_del2t(c, c->g.nrchop, f[0]->e, 1.0, f[1]->e, f[1]->ln));
      break;
    case 6: {
      const double2 *ie[9];
      for (int i = 0; i < 9; i++) ie[i] = f[i]->e;
      RET(launch_adj_columns(c, c->g.nrchop, ie, f[9]->e, f[9]->ln, f[10]->e, f[10]->ln, f[11]->e));
      break;
    }
    case 7: {
      double2 *ue[9];
      for (int i = 0; i < 9; i++) ue[i] = f[i]->e;
      const double2 *oe[3] = {f[9]->e, f[10]->e, f[11]->e};
      RET(launch_adj_products(c, ue, oe, f[12]->e));
      break;
    }
  }
  return mlg_synchronize(c);
}

// The solver's own history fields (N_psi, N_chi, N_b of the AB2 step), for
// callers that save and restore the march's state.
extern "C" int mlg_solver_history(mlg_solver *s, mlg_field **hpsi, mlg_field **hchi, mlg_field **hb) {
  if (!s || !hpsi || !hchi || !hb) return MLG_E_ARG;
  *hpsi = s->n_psi;
  *hchi = s->n_chi;
  *hb = s->n_b;
  return 0;
}
