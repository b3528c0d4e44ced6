// Adjoint kernels of the AB2-CN Boussinesq step.  The reference has no
// adjoint; these are the transposes of the operators in CALC_BOUSSI_FORCE
// (src/mod_boussinesq.f90:317-410), PROJECT (src/mod_legops.f90:2397-2648),
// PC2VEL / PC2VOR / GRAD (src/mod_legops.f90:1727-1919), DEL2
// (src/mod_legops.f90:751-820) and the energy sums of CALC_BOUSSI_ENERGY
// (src/mod_boussinesq.f90:112-232), for the inner product
// <a, b> = sum over stored entries of Re(conj(a) b), plus a%LN * b%LN for
// fields that carry a log term.
//
// Transforms need no kernels of their own: the transpose of each transform
// is the opposite transform between diagonal scalings (k_adj_scale_pff,
// k_adj_scale_pfp), because the forward transforms are a weighted Gauss sum
// (RTRAN), an FFT scaled by 1/NX (VERFFT) and a real FFT of 2 NTH samples
// scaled by 1/(2 NTH) that keeps only m < NTCHOP (HORFFT).
#include <cmath>

#include "mlegs_internal.h"

namespace mlg {
namespace {

__device__ __forceinline__ int a_nrchops(int nrchop, int m) {
  int v = nrchop - m < nrchop ? nrchop - m : nrchop;
  return v > 0 ? v : 0;
}
__device__ __forceinline__ double2 z2() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 sc(double a, double2 x) { return make_double2(a * x.x, a * x.y); }
__device__ __forceinline__ double2 pl(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 mi(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 isc(double s, double2 x) { return make_double2(-s * x.y, s * x.x); }  // (i s) x

#define A_GRID_STRIDE(n) for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (n); idx += (long)gridDim.x * blockDim.x)

inline unsigned grid_for(long n, int block) {
  long b = (n + block - 1) / block;
  return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

// PFF field a(r, m, k): rows r < NR scaled by s * f(r) and rows r >= NR
// zeroed; mode 0: f = w(rf), mode 1: f = 1 / w(rf), mode 2: f = w(r) / w(rf)
// with rf = min(r, NR-1-r) (RTRAN's Gauss sums use the weights of the first
// NRH nodes for both halves); columns m >= NTCHOP zeroed, and the m = 0
// columns too when skip_m0.
__global__ void k_adj_scale_pff(KGrid g, const double *w, double2 *a, int mode, double s, int skip_m0) {
  const long n = (long)g.ndimr * g.ntchopdim * g.nxchopdim;
  A_GRID_STRIDE(n) {
    const int r = (int)(idx % g.ndimr);
    const int mm = (int)((idx / g.ndimr) % g.ntchopdim);
    if (r >= g.nr || mm >= g.ntchop || (skip_m0 && mm == 0)) {
      a[idx] = z2();
      continue;
    }
    const int rf = r < g.nr - 1 - r ? r : g.nr - 1 - r;
    const double f = mode == 0 ? w[rf] : mode == 1 ? 1.0 / w[rf] : w[r] / w[rf];
    a[idx] = sc(s * f, a[idx]);
  }
}

// PFP field a(r, j, z), slots j < NTCHOPDIM: scaled by s0 (j = 0) or s1
// (j >= 1) for r < NR, zeroed for r >= NR
__global__ void k_adj_scale_pfp(KGrid g, double2 *a, double s0, double s1) {
  const long n = (long)g.ndimr * g.ntchopdim * g.ndimx;
  A_GRID_STRIDE(n) {
    const int r = (int)(idx % g.ndimr);
    const int j = (int)((idx / g.ndimr) % g.ntchopdim);
    const int z = (int)(idx / ((long)g.ndimr * g.ntchopdim));
    const size_t at = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    a[at] = r < g.nr ? sc(j == 0 ? s0 : s1, a[at]) : z2();
  }
}

// Transpose of CALC_BOUSSI_FORCE's pointwise terms (VPRODSUB u x omega,
// VDOTPSUB u.grad b, the Coriolis, buoyancy and stratification terms) on
// both packed theta sub-grids.  Forward, per real sample with rr = r^2:
//   o1 = a2 b3 - a3 b2 + 2 Om a2          o2 = a3 b1 - a1 b3 - 2 Om a1
//   o3 = (a1 b2 - a2 b1) / rr - w         bn = -((a1 c1 + a2 c2) / rr + a3 c3) + N^2 a3
// (a = (r u_r, r u_theta, u_z), b = (r om_r, r om_theta, om_z), c = (r db/dr,
// db/dtheta, db/dz), w = b).  In: u, om, gb (forward values), o (adjoint of
// o1..o3) and bn (adjoint of bn).  Out, in place: u <- abar, om <- bbar,
// gb <- cbar, bn <- wbar.  Entries outside r < NR, j < NTH are zeroed.
__global__ void k_adj_products(KGrid g, const double *r2, double two_omega, double n2, double2 *u0, double2 *u1, double2 *u2,
                               double2 *w0, double2 *w1, double2 *w2, double2 *c0, double2 *c1, double2 *c2,
                               const double2 *o0, const double2 *o1, const double2 *o2, double2 *bn) {
  const long n = (long)g.ndimr * g.ndimth * g.ndimx;
  A_GRID_STRIDE(n) {
    const int r = (int)(idx % g.ndimr);
    const int j = (int)((idx / g.ndimr) % g.ndimth);
    if (r >= g.nr || j >= g.nth) {
      u0[idx] = u1[idx] = u2[idx] = z2();
      w0[idx] = w1[idx] = w2[idx] = z2();
      c0[idx] = c1[idx] = c2[idx] = z2();
      bn[idx] = z2();
      continue;
    }
    const double rr = r2[r];  // the device radial table holds r^2
    const double2 A1 = u0[idx], A2 = u1[idx], A3 = u2[idx];
    const double2 B1 = w0[idx], B2 = w1[idx], B3 = w2[idx];
    const double2 C1 = c0[idx], C2 = c1[idx], C3 = c2[idx];
    const double2 O1 = o0[idx], O2 = o1[idx], O3 = o2[idx], BN = bn[idx];
    double2 r_[13];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      auto G = [h](double2 v) { return h ? v.y : v.x; };
      const double a1 = G(A1), a2 = G(A2), a3 = G(A3), b1 = G(B1), b2 = G(B2), b3 = G(B3);
      const double c1 = G(C1), c2 = G(C2), c3 = G(C3), o1 = G(O1), o2 = G(O2), o3 = G(O3), bb = G(BN);
      double v[13];
      v[0] = -b3 * o2 + (b2 / rr) * o3 - (c1 / rr) * bb - two_omega * o2;
      v[1] = b3 * o1 - (b1 / rr) * o3 - (c2 / rr) * bb + two_omega * o1;
      v[2] = -b2 * o1 + b1 * o2 - c3 * bb + n2 * bb;
      v[3] = a3 * o2 - (a2 / rr) * o3;
      v[4] = -a3 * o1 + (a1 / rr) * o3;
      v[5] = a2 * o1 - a1 * o2;
      v[6] = -(a1 / rr) * bb;
      v[7] = -(a2 / rr) * bb;
      v[8] = -a3 * bb;
      v[9] = -o3;
#pragma unroll
      for (int q = 0; q < 10; q++) {
        if (h == 0) r_[q].x = v[q];
        else r_[q].y = v[q];
      }
    }
    u0[idx] = r_[0], u1[idx] = r_[1], u2[idx] = r_[2];
    w0[idx] = r_[3], w1[idx] = r_[4], w2[idx] = r_[5];
    c0[idx] = r_[6], c1[idx] = r_[7], c2[idx] = r_[8];
    bn[idx] = r_[9];
  }
}

// Transpose of PROJECT's products (src/mod_legops.f90:2496-2648).  Forward,
// per (m, k), with EO(A, f)(n) = sum_j A(n, j) (f(j) +- f(NR-1-j)) (+ for
// even n - m, - for odd) and OE the same with the signs swapped:
//   psi  = -i m EO(V, ru_r) - OE(D, ru_theta)
//   chi2 = i k OE(D, ru_r) + m k EO(V, ru_theta) - EO(T, u_z)
// In: psi, chi2 adjoints (FFF, chopped); out: the three PFF adjoints (all
// rows written).  V, D, T: vdt[(t NTCHOPDIM + m) NRCHOPDIM + n][j], j < NRH.
// The psi adjoint is read through the transposes of CHOPDO and of PROJECT's
// fix at infinity of m = 0 (psibar(n) - AT1(n)/AT1(0) psibar(0)), chi2's
// through CHOPDO's.  Per (m, k) and node pair (j, NR-1-j), with the
// coefficient factors a1 = i m psibar, a2 = -i k chi2bar, a3 = -psibar,
// a4 = m k chi2bar, a5 = -chi2bar and S_e / S_o sums over even / odd n - m:
//   ru_r:  row j  Ve a1 + Vo a1 + De a2 + Do a2,   row NR-1-j  (Ve - Vo) a1 + (Do - De) a2
//   ru_t:  row j  De a3 + Do a3 + Ve a4 + Vo a4,   row NR-1-j  (Do - De) a3 + (Ve - Vo) a4
//   u_z:   row j  Te a5 + To a5,                   row NR-1-j  (Te - To) a5
// A block takes one m and ADJ_PKB consecutive k: the factors go to LDS, each
// thread a node j, and a table row is loaded once for the ADJ_PKB columns.
// One column per block: at 128x64x64 that is 946 one-wave blocks (28 us),
// where four columns per block left 242 (70 us).
#define ADJ_PKB 1
__global__ void __launch_bounds__(64) k_adj_project(KGrid g, const double *vdt, const double *ak, const double *at1,
                                                    const double2 *psib, const double2 *chib, double2 *vr, double2 *vp,
                                                    double2 *vz) {
  extern __shared__ __align__(16) double2 fs[];  // [ADJ_PKB][5][nn]
  const int mm = blockIdx.x, k0 = blockIdx.y * ADJ_PKB;
  const int nn = mm < g.ntchop ? a_nrchops(g.nrchop, mm * g.minc) : 0;
  const double mv = (double)(mm * g.minc);
  const size_t tstride = (size_t)g.ntchopdim * g.nrchopdim * g.nrh;
  const int nk = g.nxchopdim - k0 < ADJ_PKB ? g.nxchopdim - k0 : ADJ_PKB;
  for (int i = threadIdx.x; i < ADJ_PKB * nn; i += blockDim.x) {
    const int q = i / (nn > 0 ? nn : 1), ii = i - q * nn;
    double2 P = z2(), X = z2();
    double kv = 0.0;
    if (q < nk) {
      const size_t fcol = (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * (k0 + q));
      P = mm == 0 ? mi(psib[fcol + ii], sc(at1[ii] / at1[0], psib[fcol])) : psib[fcol + ii];
      X = chib[fcol + ii];
      kv = ak[mm + (size_t)g.ntchopdim * (k0 + q)];
    }
    double2 *f = fs + (size_t)q * 5 * nn;
    f[ii] = isc(mv, P);
    f[nn + ii] = isc(-kv, X);
    f[2 * nn + ii] = sc(-1.0, P);
    f[3 * nn + ii] = sc(mv * kv, X);
    f[4 * nn + ii] = sc(-1.0, X);
  }
  __syncthreads();
  for (int q = 0; q < nk; q++) {
    const size_t pcol = (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * (k0 + q));
    for (int r = g.nr + threadIdx.x; r < g.ndimr; r += blockDim.x) vr[pcol + r] = vp[pcol + r] = vz[pcol + r] = z2();
  }
  for (int j = threadIdx.x; j < g.nrh; j += blockDim.x) {
    // [column][e/o] sums: r1 = V a1, r2 = D a2, p3 = D a3, p4 = V a4, z5 = T a5
    double2 r1[ADJ_PKB][2], r2[ADJ_PKB][2], p3[ADJ_PKB][2], p4[ADJ_PKB][2], z5[ADJ_PKB][2];
#pragma unroll
    for (int q = 0; q < ADJ_PKB; q++)
#pragma unroll
      for (int e = 0; e < 2; e++) r1[q][e] = r2[q][e] = p3[q][e] = p4[q][e] = z5[q][e] = z2();
    // rows in pairs (even n - m, then odd), so the parity index is static
    for (int i2 = 0; i2 < nn; i2 += 2) {
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int ii = i2 + e;
        if (ii >= nn) break;
        const size_t tb = ((size_t)mm * g.nrchopdim + ii) * g.nrh + j;
        const double V = vdt[tb], D = vdt[tstride + tb], T = vdt[2 * tstride + tb];
#pragma unroll
        for (int q = 0; q < ADJ_PKB; q++) {
          const double2 *f = fs + (size_t)q * 5 * nn;
          r1[q][e] = pl(r1[q][e], sc(V, f[ii]));
          r2[q][e] = pl(r2[q][e], sc(D, f[nn + ii]));
          p3[q][e] = pl(p3[q][e], sc(D, f[2 * nn + ii]));
          p4[q][e] = pl(p4[q][e], sc(V, f[3 * nn + ii]));
          z5[q][e] = pl(z5[q][e], sc(T, f[4 * nn + ii]));
        }
      }
    }
#pragma unroll
    for (int q = 0; q < ADJ_PKB; q++) {
      if (q >= nk) break;
      const size_t pcol = (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * (k0 + q));
      vr[pcol + j] = pl(pl(r1[q][0], r1[q][1]), pl(r2[q][0], r2[q][1]));
      vr[pcol + g.nr - 1 - j] = pl(mi(r1[q][0], r1[q][1]), mi(r2[q][1], r2[q][0]));
      vp[pcol + j] = pl(pl(p3[q][0], p3[q][1]), pl(p4[q][0], p4[q][1]));
      vp[pcol + g.nr - 1 - j] = pl(mi(p3[q][1], p3[q][0]), mi(p4[q][0], p4[q][1]));
      vz[pcol + j] = pl(z5[q][0], z5[q][1]);
      vz[pcol + g.nr - 1 - j] = mi(z5[q][0], z5[q][1]);
    }
  }
}

// Transpose of VEL2PLN (src/mod_legops.f90:2277-2339) on the (m = 0,
// k = 0) column of ru_theta in PFF.  Forward: LN = -s/2 with
// s = sum_j sigma(j) Re ru(j), then ru(j) += LN c(j) (the vectors c, sigma
// are cs[0..NR-1], cs[NR..2NR-1], built on the host for the truncation).
// Adjoint: Re rubar(j) -= sigma(j)/2 * sum_l c(l) Re rubar(l).  One block.
__global__ void __launch_bounds__(256) k_adj_vel2pln(KGrid g, const double *cs, double2 *vp) {
  __shared__ double red[256];
  double part = 0.0;
  for (int i = threadIdx.x; i < g.nr; i += blockDim.x) part += cs[i] * vp[i].x;
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double cdot = red[0];
  for (int i = threadIdx.x; i < g.nr; i += blockDim.x) vp[i].x -= 0.5 * cs[g.nr + i] * cdot;
}

// Transpose of the value-at-infinity fix-up of m = 0 (CALCAT1,
// src/mod_scalar3.f90:286-310): forward a(0) <- a(0) - sum_n AT1(n) a(n) /
// AT1(0) over n < NRCHOP; adjoint abar(n) -= AT1(n)/AT1(0) abar(0) (which
// leaves abar(0) = 0).  One block per k.
__global__ void __launch_bounds__(64) k_adj_fixinf(KGrid g, const double *at1, double2 *a) {
  double2 *col = a + (size_t)g.nrchopdim * ((size_t)g.ntchopdim * blockIdx.x);
  const double2 a0 = col[0];
  __syncthreads();
  for (int n = threadIdx.x; n < g.nrchop; n += blockDim.x) col[n] = mi(col[n], sc(at1[n] / at1[0], a0));
}

// Transpose of PC2VEL, PC2VOR and GRAD (src/mod_legops.f90:1727-1919) at the
// truncation NRCHOP of g (NRCHOP+3 inside CALC_BOUSSI_FORCE), accumulated
// into the adjoints of psi, chi and b rows n < NRCHOPS(m) at nrchop_out.
// Forward per (m, k), X = XXDX, M = MULXM, Dh = DELSQH, L = del^2_h band:
//   ru_r = i m psi + i k X chi          ru_t = -X psi - m k chi
//   u_z = -M M Dh chi                   oz' = L chi - k^2 chi
//   ro_r = -i m oz' + i k X psi         ro_t = X oz' - k m psi
//   o_z = -M M Dh psi
//   gb_r = X b, gb_t = i m b, gb_z = i k b
// plus the log terms at (m, k) = (0, 0): X adds LN/N(n) to rows 0, 1, Dh
// sets row 0 to LN/(L^2 N(0)), DEL2 adds LN d(n) to rows 0..2.  in[9]:
// adjoints of ru_r ru_t u_z ro_r ro_t o_z gb_r gb_t gb_z (FFF at g).
struct AdjColIn {
  const double2 *e[9];
};
__global__ void __launch_bounds__(64) k_adj_columns(KGrid g, int nrchop_out, const double *bxxdx, const double *bxm,
                                                    const double *bdel2h, const double *ak, AdjColIn in, double2 *psib,
                                                    double *psib_ln, double2 *chib, double *chib_ln, double2 *bb) {
  extern __shared__ __align__(16) double2 cs[];
  const int mm = blockIdx.x, k = blockIdx.y;
  if (mm >= g.ntchop) return;
  const int nn = a_nrchops(g.nrchop, mm * g.minc), no = a_nrchops(nrchop_out, mm * g.minc);
  if (nn < 1) return;
  const int L = nn + 4;  // rows with two zero guards on each side
  double2 *in_s = cs;         // [9][L]
  double2 *t_s = cs + 9 * L;  // [6][L]: Ypsi, Ychi, Ozt, T1, T2, gb
  const double mv = (double)(mm * g.minc), kv = ak[mm + (size_t)g.ntchopdim * k];
  const size_t col = (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
  const double *Xb = bxxdx + (size_t)mm * g.nband * 3, *Mb = bxm + (size_t)mm * g.nband * 3, *Lb = bdel2h + (size_t)mm * g.nband * 5;
  for (int i = threadIdx.x; i < 9 * L; i += blockDim.x) {
    const int f = i / L, rr = i % L - 2;
    in_s[i] = (rr >= 0 && rr < nn) ? in.e[f][col + rr] : z2();
  }
  for (int i = threadIdx.x; i < 6 * L; i += blockDim.x) t_s[i] = z2();
  __syncthreads();
  auto IN = [&](int f, int r) -> double2 { return in_s[f * L + r + 2]; };
  auto T = [&](int f, int r) -> double2 & { return t_s[f * L + r + 2]; };
  // transposed tri-band of table B at column c over rows [0, nn)
  auto tri_t = [&](const double *B, int f, int c, bool tin) -> double2 {
    double2 s = z2();
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int row = c - j + 1;
      if (row >= 0 && row < nn) s = pl(s, sc(B[(size_t)row * 3 + j], tin ? T(f, row) : IN(f, row)));
    }
    return s;
  };
  // stage 1: Ypsi = -Rut - i k Ror (into X^T), Ychi = -i k Rur, Ozt = i m Ror + X^T Rot, T1 = M^T Oz, T2 = M^T Uz
  for (int c = threadIdx.x; c < nn; c += blockDim.x) {
    T(0, c) = mi(sc(-1.0, IN(1, c)), isc(kv, IN(3, c)));
    T(1, c) = isc(-kv, IN(0, c));
    T(2, c) = pl(isc(mv, IN(3, c)), tri_t(Xb, 4, c, false));
    T(3, c) = tri_t(Mb, 5, c, false);
    T(4, c) = tri_t(Mb, 2, c, false);
  }
  __syncthreads();
  // stage 2: per row, the outputs
  for (int c = threadIdx.x; c < no; c += blockDim.x) {
    const int n = (int)mv + c;
    const double dh = -(double)n * (n + 1) / g.ell2;
    const double2 xps = tri_t(Xb, 0, c, true), xch = tri_t(Xb, 1, c, true);
    const double2 m1 = tri_t(Mb, 3, c, true), m2 = tri_t(Mb, 4, c, true);
    double2 lo = mi(sc(-kv * kv, T(2, c)), z2());
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const int row = c - j + 2;
      if (row >= 0 && row < nn) lo = pl(lo, sc(Lb[(size_t)row * 5 + j], T(2, row)));
    }
    double2 p = isc(-mv, IN(0, c));
    p = pl(p, xps);
    p = mi(p, sc(kv * mv, IN(4, c)));
    p = mi(p, sc(dh, m1));
    double2 h = pl(lo, sc(-mv * kv, IN(1, c)));
    h = pl(h, xch);
    h = mi(h, sc(dh, m2));
    double2 bv = tri_t(Xb, 6, c, false);
    bv = pl(bv, isc(-mv, IN(7, c)));
    bv = pl(bv, isc(-kv, IN(8, c)));
    psib[col + c] = pl(psib[col + c], p);
    chib[col + c] = pl(chib[col + c], h);
    bb[col + c] = pl(bb[col + c], bv);
  }
  if (mm == 0 && k == 0 && threadIdx.x == 0) {
    // log terms: rut = -X psi (LN/N rows 0, 1), o_z = -M M Dh psi, u_z = -M M Dh chi, DEL2's rows 0..2
    const double m1 = tri_t(Mb, 3, 0, true).x, m2 = tri_t(Mb, 4, 0, true).x;
    double lp = -(IN(1, 0).x / g.norm0 + (nn > 1 ? IN(1, 1).x / g.norm1 : 0.0));
    lp += -m1 / g.ell2 / g.norm0;
    double lc = -m2 / g.ell2 / g.norm0;
    lc += T(2, 0).x * (4.0 / 3 / g.ell2 / g.norm0);
    if (nn > 1) lc += T(2, 1).x * (-2.0 / g.ell2 / g.norm1);
    if (nn > 2) lc += T(2, 2).x * (2.0 / 3 / g.ell2 / g.norm2);
    psib_ln[0] += lp;
    chib_ln[0] += lc;
  }
}

// out += s (L^T y - k^2 y) on rows n < NRCHOPS(m) at nrchop_out, L the del^2_h
// band at the truncation of g (transpose of DEL2's band part); ln_out +=
// s * (DEL2's log-term rows of y at (0, 0)) when ln_out is given.
__global__ void k_adj_del2t(KGrid g, int nrchop_out, const double *bdel2h, const double *ak, const double2 *y, double s,
                            double2 *out, double *ln_out) {
  const long n = (long)g.nrchopdim * g.ntchopdim * g.nxchopdim;
  A_GRID_STRIDE(n) {
    const int c = (int)(idx % g.nrchopdim);
    const int mm = (int)((idx / g.nrchopdim) % g.ntchopdim);
    const int k = (int)(idx / ((long)g.nrchopdim * g.ntchopdim));
    if (mm >= g.ntchop) continue;
    const int nn = a_nrchops(g.nrchop, mm * g.minc), no = a_nrchops(nrchop_out, mm * g.minc);
    const size_t col = (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
    if (c == 0 && mm == 0 && k == 0 && ln_out) {
      double v = y[col].x * (4.0 / 3 / g.ell2 / g.norm0);
      if (nn > 1) v += y[col + 1].x * (-2.0 / g.ell2 / g.norm1);
      if (nn > 2) v += y[col + 2].x * (2.0 / 3 / g.ell2 / g.norm2);
      ln_out[0] += s * v;
    }
    if (c >= no || c >= nn) continue;
    const double kv = ak[mm + (size_t)g.ntchopdim * k];
    const double *Lb = bdel2h + (size_t)mm * g.nband * 5;
    double2 t = sc(-kv * kv, y[col + c]);
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const int row = c - j + 2;
      if (row >= 0 && row < nn) t = pl(t, sc(Lb[(size_t)row * 5 + j], y[col + row]));
    }
    out[col + c] = pl(out[col + c], sc(s, t));
  }
}

// Transpose of the AB2 update and CN viscosity of STEP_BOUSSI_AB_CN for up
// to three fields in one launch (grid row = field).  Forward, per field:
//   u* = u + dt (1.5 N - 0.5 N_old);  N_old' = N;
//   u' = IHELM(-DEL2(u) - ALP u*, ALP) (DEL2 at NRCHOP+2, chopped to NRCHOP), or u' = u* without CN.
// In: wk = IHELM^T(u' bar) (CN fields), ub = u' bar (others), hb = N_old' bar.
// Out: ub = u bar = u* bar - DEL2^T(wk), nb = N bar = CHOPDO(hb) + 1.5 dt u* bar,
// hb = N_old bar = -0.5 dt u* bar, with u* bar = -ALP wk (CN) or u' bar.
// g carries the truncation NRCHOP+2; rows are kept below NRCHOPS at nrchop_out.
struct AdjCnJobs {
  AdjCnField f[3];
};
__global__ void k_adj_cn(KGrid g, int nrchop_out, const double *bdel2h, const double *ak, AdjCnJobs jobs, double dt) {
  const AdjCnField F = jobs.f[blockIdx.y];
  const long n = (long)g.nrchopdim * g.ntchopdim * g.nxchopdim;
  A_GRID_STRIDE(n) {
    const int c = (int)(idx % g.nrchopdim);
    const int mm = (int)((idx / g.nrchopdim) % g.ntchopdim);
    const int k = (int)(idx / ((long)g.nrchopdim * g.ntchopdim));
    const int no = mm < g.ntchop ? a_nrchops(nrchop_out, mm * g.minc) : 0;
    const bool in = c < no;
    double2 st = z2(), u = z2();
    if (in) {
      if (F.cn) {
        const int nn = a_nrchops(g.nrchop, mm * g.minc);
        const size_t col = idx - c;
        const double kv = ak[mm + (size_t)g.ntchopdim * k];
        const double *Lb = bdel2h + (size_t)mm * g.nband * 5;
        double2 t = sc(-kv * kv, F.wk[idx]);
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const int row = c - j + 2;
          if (row >= 0 && row < nn) t = pl(t, sc(Lb[(size_t)row * 5 + j], F.wk[col + row]));
        }
        st = sc(-F.alp, F.wk[idx]);
        u = mi(st, t);
      } else {
        st = F.ub[idx];
        u = st;
      }
    }
    const double2 h = in ? F.hb[idx] : z2();
    F.nb[idx] = pl(h, sc(1.5 * dt, st));
    F.hb[idx] = sc(-0.5 * dt, st);
    F.ub[idx] = u;
    if (idx == 0) {
      if (F.ub_ln) {
        const int nn = a_nrchops(g.nrchop, 0);
        double stl = F.cn ? -F.alp * F.wk_ln[0] : F.ub_ln[0], ul = stl;
        if (F.cn) {
          ul -= F.wk[0].x * (4.0 / 3 / g.ell2 / g.norm0);
          if (nn > 1) ul -= F.wk[1].x * (-2.0 / g.ell2 / g.norm1);
          if (nn > 2) ul -= F.wk[2].x * (2.0 / 3 / g.ell2 / g.norm2);
        }
        F.nb_ln[0] = F.hb_ln[0] + 1.5 * dt * stl;
        F.hb_ln[0] = -0.5 * dt * stl;
        F.ub_ln[0] = ul;
      }
    }
  }
}

// y = a y + b1 x1 + b2 x2 (x1, x2 may be null), n entries
__global__ void k_adj_lincomb(double2 *y, double a, const double2 *x1, double b1, const double2 *x2, double b2, long n) {
  A_GRID_STRIDE(n) {
    double2 v = sc(a, y[idx]);
    if (x1) v = pl(v, sc(b1, x1[idx]));
    if (x2) v = pl(v, sc(b2, x2[idx]));
    y[idx] = v;
  }
}
__global__ void k_adj_ln_lincomb(double *y, double a, const double *x1, double b1, const double *x2, double b2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double v = a * y[0];
    if (x1) v += b1 * x1[0];
    if (x2) v += b2 * x2[0];
    y[0] = v;
  }
}

// ---------------------------------------------------------------- launchers
int launch_adj_scale_pff(mlg_ctx *c, double2 *a, int mode, double s, int skip_m0) {
  KGrid g = kgrid(c);
  const long n = (long)g.ndimr * g.ntchopdim * g.nxchopdim;
  hipLaunchKernelGGL(k_adj_scale_pff, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, g, c->d.w, a, mode, s, skip_m0);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_scale_pfp(mlg_ctx *c, double2 *a, double s0, double s1) {
  KGrid g = kgrid(c);
  const long n = (long)g.ndimr * g.ntchopdim * g.ndimx;
  hipLaunchKernelGGL(k_adj_scale_pfp, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, g, a, s0, s1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_products(mlg_ctx *c, double2 *const u[9], const double2 *const o[3], double2 *bn) {
  KGrid g = kgrid(c);
  const long n = (long)g.ndimr * g.ndimth * g.ndimx;
  hipLaunchKernelGGL(k_adj_products, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, g, c->d.r, 2 * c->g.p.omega,
                     c->g.p.bv0 * c->g.p.bv0, u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8], o[0], o[1], o[2], bn);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_project(mlg_ctx *c, const double2 *psib, const double2 *chib, double2 *vr, double2 *vp, double2 *vz) {
  KGrid g = kgrid(c);
  const size_t lds = sizeof(double2) * 5 * ADJ_PKB * (size_t)(g.nrchop > 0 ? g.nrchop : 1);
  hipLaunchKernelGGL(k_adj_project, dim3(g.ntchopdim, (g.nxchopdim + ADJ_PKB - 1) / ADJ_PKB), dim3(64), lds, c->stream, g,
                     c->d.vdt, c->d.ak, c->d.at1, psib, chib, vr, vp, vz);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
// VEL2PLN's vectors c and sigma at the current truncation (cached per
// truncation): for i < NRH,
//   c(i) = PF(i,0)/N(0) + PF(i,1)/N(1),   c(NR-1-i) = PF(i,0)/N(0) - PF(i,1)/N(1),
//   sigma(i) = w(i) sum_n e^{LOGNORM(n)} PF(i,n),
//   sigma(NR-1-i) = w(i) sum_n (-1)^n e^{LOGNORM(n)} PF(i,n),   n < NRCHOPS(0)
static const double *vel2pln_vectors(mlg_ctx *c, int *rc) {
  const Grid &G = c->g;
  for (auto &v : c->adjv)
    if (v.first == G.nrchop) return v.second;
  const int nr = G.nr, nrh = G.nrh, nn = nrchops_of(G.nrchop, 0);
  std::vector<double> h(2 * (size_t)nr, 0.0);
  const double n0 = G.norm[0], n1 = G.norm[1];
  for (int i = 0; i < nrh; i++) {
    const double p0 = G.pf[i], p1 = G.pf[(size_t)nrh + i];
    h[i] = p0 / n0 + p1 / n1;
    h[nr - 1 - i] = p0 / n0 - p1 / n1;
    double se = 0.0, so = 0.0;
    for (int n = 0; n < nn; n++) {
      const double t = std::exp(G.lognorm[n]) * G.pf[(size_t)n * nrh + i];
      se += t;
      so += (n & 1) ? -t : t;
    }
    h[nr + i] = G.w[i] * se;
    h[nr + nr - 1 - i] = G.w[i] * so;
  }
  double *d = nullptr;
  if (hipMalloc((void **)&d, sizeof(double) * h.size()) != hipSuccess ||
      hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    *rc = MLG_E_HIP;
    return nullptr;
  }
  c->adjv.push_back({G.nrchop, d});
  return d;
}

int launch_adj_vel2pln(mlg_ctx *c, double2 *vp) {
  KGrid g = kgrid(c);
  if (g.ntchop < 1 || g.nrchop < 1) return 0;
  int rc = 0;
  const double *cs = vel2pln_vectors(c, &rc);
  if (!cs) return rc;
  hipLaunchKernelGGL(k_adj_vel2pln, dim3(1), dim3(256), 0, c->stream, g, cs, vp);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_fixinf(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  if (g.nrchop < 1) return 0;
  hipLaunchKernelGGL(k_adj_fixinf, dim3(g.nxchopdim), dim3(64), 0, c->stream, g, c->d.at1, a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_columns(mlg_ctx *c, int nrchop_out, const double2 *const in[9], double2 *psib, double *psib_ln, double2 *chib,
                       double *chib_ln, double2 *bb) {
  KGrid g = kgrid(c);
  AdjColIn io;
  for (int i = 0; i < 9; i++) io.e[i] = in[i];
  const size_t lds = sizeof(double2) * 15 * (size_t)(g.nrchopdim + 4);
  hipLaunchKernelGGL(k_adj_columns, dim3(g.ntchopdim, g.nxchopdim), dim3(64), lds, c->stream, g, nrchop_out, c->d.bxxdx,
                     c->d.bxm, c->d.bdel2h, c->d.ak, io, psib, psib_ln, chib, chib_ln, bb);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_del2t(mlg_ctx *c, int nrchop_out, const double2 *y, double s, double2 *out, double *ln_out) {
  KGrid g = kgrid(c);
  const long n = (long)g.nrchopdim * g.ntchopdim * g.nxchopdim;
  hipLaunchKernelGGL(k_adj_del2t, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, g, nrchop_out, c->d.bdel2h, c->d.ak, y, s,
                     out, ln_out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_cn(mlg_ctx *c, int nrchop_out, int nf, const AdjCnField *f, double dt) {
  KGrid g = kgrid(c);
  AdjCnJobs jobs;
  for (int i = 0; i < nf; i++) jobs.f[i] = f[i];
  const long n = (long)g.nrchopdim * g.ntchopdim * g.nxchopdim;
  hipLaunchKernelGGL(k_adj_cn, dim3(grid_for(n, 256), nf), dim3(256), 0, c->stream, g, nrchop_out, c->d.bdel2h, c->d.ak, jobs,
                     dt);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_lincomb(mlg_ctx *c, double2 *y, double a, const double2 *x1, double b1, const double2 *x2, double b2, long n) {
  hipLaunchKernelGGL(k_adj_lincomb, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, y, a, x1, b1, x2, b2, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_adj_ln_lincomb(mlg_ctx *c, double *y, double a, const double *x1, double b1, const double *x2, double b2) {
  hipLaunchKernelGGL(k_adj_ln_lincomb, dim3(1), dim3(64), 0, c->stream, y, a, x1, b1, x2, b2);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
