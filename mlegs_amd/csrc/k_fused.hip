// Nonlinear terms in physical space without physical-space fields in HBM.
//
// CALC_BOUSSI_FORCE / CALC_BOUSSI_NONLIN (src/mod_boussinesq.f90:235-410)
// take nine fields to physical space (HORFFT backward), form u x omega and
// u . grad b there (VPRODSUB/VDOTPSUB, src/mod_legops.f90:1972-2232, plus
// FORCE's linear terms) and bring four results back (HORFFT forward).  This
// kernel does all of that for two radial pencils of one z plane at a time,
// reading the inputs in PFP (r, m, z) and writing the outputs in PFP: the
// 512-point physical theta lines live only in LDS.
//
// Block: 256 threads in groups of 32 (2 pencils x 16 threads at NTH = 256);
// a group transforms one field with the four-step FFT of k_fft.hip.  LDS:
// eight slots of 2 x NTH complex values (64 KiB at NTH = 256).
//   phase 1: inverse transforms of rur, rup, uz, ror, rop, oz, rbr, rbp
//   products A (every point): X = rup oz - uz rop (+ 2 Omega rup),
//            Y = uz ror - rur oz (- 2 Omega rur), Z = (rur rop - rup ror) / r^2,
//            T = rur rbr + rup rbp
//   phase 2: forward transforms of X, Y (and Z without linear terms); then
//            inverse transforms of bz (and of W = b for FORCE)
//   products B: N_b = -(T / r^2 + uz bz) (+ Nbar^2 uz); Z = Z - W (FORCE)
//   phase 3: forward transforms of N_b (and Z for FORCE)
// Two pencils per block keep the LDS within two blocks per CU; the price is
// 32-byte runs in the PFP reads and writes, which the XCD-aware block order
// turns into shared L2 lines.
// Every value goes through the operations of the separate HORFFT and
// product kernels, in the same order.  NTH = 256 has its own kernel (below)
// with the same products and an 8 x 8 x 4 factorisation of the transforms.
#include "mlegs_internal.h"
#include "xcd.h"

namespace mlg {

// The theta transforms round like any FFT (not operation for operation like
// a reference library): products and sums may contract to FMAs here.
#pragma clang fp contract(fast)

__device__ __forceinline__ double2 fz_cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 fz_conj(double2 a) { return make_double2(a.x, -a.y); }

template <int R>
struct FLog2 {
  static constexpr int v = R <= 1 ? 0 : 1 + FLog2<R / 2>::v;
};
template <>
struct FLog2<1> {
  static constexpr int v = 0;
};
__device__ __forceinline__ constexpr int fz_bitrev(int i, int lg) {
  int r = 0;
  for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
  return r;
}
template <int R>
__device__ __forceinline__ void fz_dft(double2 (&v)[R], const double2 (&wR)[R / 2 > 0 ? R / 2 : 1]) {
  constexpr int LG = FLog2<R>::v;
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int j = fz_bitrev(i, LG);
    if (j > i) {
      double2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int half = 1; half < R; half *= 2) {
#pragma unroll
    for (int g0 = 0; g0 < R; g0 += 2 * half) {
#pragma unroll
      for (int p = 0; p < half; p++) {
        const double2 w = wR[p * (R / (2 * half))];
        double2 u = v[g0 + p], x = fz_cmul(v[g0 + p + half], w);
        v[g0 + p] = make_double2(u.x + x.x, u.y + x.y);
        v[g0 + p + half] = make_double2(u.x - x.x, u.y - x.y);
      }
    }
  }
}
template <int R>
__device__ __forceinline__ void fz_wr(double2 (&wR)[R / 2 > 0 ? R / 2 : 1], const double2 *__restrict__ tw, int N, bool inv) {
#pragma unroll
  for (int j = 0; j < (R / 2 > 0 ? R / 2 : 1); j++) {
    double2 w = (R > 1) ? tw[j * (N / R)] : make_double2(1.0, 0.0);
    if (inv) w.y = -w.y;
    wR[j] = w;
  }
}

constexpr int FP = 2;  // pencils per block

// Four-step transform of the two pencils held by a group.  Pass 1 input in
// v1 (thread t < R2 holds x[R2 n1 + t]); the exchange goes through `s`
// (slot layout [index][pencil]); on return thread t < R1 holds
// X[t + R1 k2] in v2.  All threads of the block call it (barriers).
template <int R1, int R2>
__device__ __forceinline__ void fz_fft(bool act, double2 (&v1)[R1], double2 (&v2)[R2], double2 *s, int p, int t,
                                       const double2 *__restrict__ tw, bool inv) {
  constexpr int N = R1 * R2;
  if (act && t < R2) {
    double2 w1[R1 / 2 > 0 ? R1 / 2 : 1];
    fz_wr<R1>(w1, tw, N, inv);
    fz_dft<R1>(v1, w1);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      double2 w = tw[t * k1];
      if (inv) w.y = -w.y;
      s[(k1 * R2 + t) * FP + p] = (k1 == 0 || t == 0) ? v1[k1] : fz_cmul(v1[k1], w);
    }
  }
  __syncthreads();
  if (act && t < R1) {
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) v2[n2] = s[(t * R2 + n2) * FP + p];
    double2 w2[R2 / 2 > 0 ? R2 / 2 : 1];
    fz_wr<R2>(w2, tw, N, inv);
    fz_dft<R2>(v2, w2);
  }
  __syncthreads();
}

// HORFFT backward of one PFP field into physical slots of `s`.
template <int R1, int R2>
__device__ void fz_inverse(bool act, const KGrid &g, const double2 *__restrict__ src, double2 *s, int r0, int z, int p, int t,
                           const double2 *__restrict__ tw, const double2 *__restrict__ twh) {
  constexpr int N = R1 * R2;
  double2 v1[R1], v2[R2];
  const int rr = r0 + p;
  const int mmax = g.ntchopdim;
  if (act && t < R2) {
    auto X = [&](int m) -> double2 {
      double2 v = make_double2(0.0, 0.0);
      if (rr < g.nr && m < g.ntchop) v = src[(size_t)rr + (size_t)g.ndimr * ((size_t)m + (size_t)g.ndimth * z)];
      if (m == 0) v.y = 0.0;
      return v;
    };
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
      const int j = R2 * n1 + t;
      double2 xm = (j < mmax) ? X(j) : make_double2(0.0, 0.0);
      const int jc = N - j;
      double2 xc = (jc < mmax && jc > 0 && j > 0) ? fz_conj(X(jc)) : make_double2(0.0, 0.0);
      double2 e = make_double2(xm.x + xc.x, xm.y + xc.y);
      double2 d = make_double2(xm.x - xc.x, xm.y - xc.y);
      double2 o = fz_cmul(d, fz_conj(twh[j]));
      v1[n1] = make_double2(e.x - o.y, e.y + o.x);
    }
  }
  fz_fft<R1, R2>(act, v1, v2, s, p, t, tw, true);
  if (act && t < R1) {
    const bool live = rr < g.nr;
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) s[(t + R1 * k2) * FP + p] = live ? v2[k2] : make_double2(0.0, 0.0);
  }
  __syncthreads();
}

// HORFFT forward of physical slots in `s` into one PFP field.
template <int R1, int R2>
__device__ void fz_forward(bool act, const KGrid &g, double2 *s, double2 *__restrict__ dst, int r0, int z, int p, int t,
                           int gt, const double2 *__restrict__ tw, const double2 *__restrict__ twh) {
  constexpr int N = R1 * R2;
  double2 v1[R1], v2[R2];
  if (act && t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) v1[n1] = s[(R2 * n1 + t) * FP + p];
  }
  __syncthreads();
  fz_fft<R1, R2>(act, v1, v2, s, p, t, tw, false);
  if (act && t < R1) {
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) s[(t + R1 * k2) * FP + p] = v2[k2];
  }
  __syncthreads();
  if (act) {
    const double den = (double)(2 * N);
    const int mmax = g.ntchopdim;
    for (int idx = gt; idx < FP * mmax; idx += FP * (R1 > R2 ? R1 : R2)) {
      const int pr = idx % FP, mm = idx / FP;
      const int rr = r0 + pr;
      if (rr >= g.ndimr) continue;
      double2 zm = s[(mm % N) * FP + pr], zc = fz_conj(s[((N - mm) % N) * FP + pr]);
      double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
      double2 wd = fz_cmul(twh[mm], d);
      double2 x = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
      x = make_double2(x.x / den, x.y / den);
      if (mm == 0) x.y = 0.0;
      if (mm >= g.ntchop || rr >= g.nr) x = make_double2(0.0, 0.0);
      dst[(size_t)rr + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = x;
    }
  }
}

struct FusedIO {
  const double2 *in[10];  // rur rup uz ror rop oz rbr rbp bz [W]
  double2 *out[4];        // r N_r, r N_theta, N_z, N_b (PFP)
  double *zl;             // an LN coefficient the launch sets to zero (N_b's), or null
};

template <int R1, int R2>
__global__ void __launch_bounds__(256) k_fused_products(KGrid g, const double *__restrict__ rsq, const double2 *__restrict__ tw,
                                                        const double2 *__restrict__ twh, FusedIO io, double two_omega,
                                                        double n2, int linear) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && io.zl) io.zl[0] = 0.0;
  constexpr int N = R1 * R2, TPP = R1 > R2 ? R1 : R2, GS = FP * TPP, NG = 256 / GS, SL = FP * N;
  static_assert(NG >= 8, "eight concurrent transforms");
  __shared__ double2 S[8 * SL];
  // logical block = (pencil pair, z) with the pair fastest: the four pairs
  // of one 128-byte line run on one XCD
  const int nrp = (g.ndimr + FP - 1) / FP;
  const int lid = xcd_logical(nrp * g.ndimx);
  if (lid < 0) return;
  const int z = lid / nrp, r0 = (lid % nrp) * FP;
  const int grp = threadIdx.x / GS, gt = threadIdx.x % GS, p = gt % FP, t = gt / FP;
  auto slot = [&](int q) { return S + (size_t)q * SL; };

  // phase 1: rur rup uz ror rop oz rbr rbp -> slots 0..7
  {
    const bool act = grp < 8;
    const int q = act ? grp : 0;
    fz_inverse<R1, R2>(act, g, io.in[q], slot(q), r0, z, p, t, tw, twh);
  }
  // products A
  for (int q = threadIdx.x; q < SL; q += 256) {
    const int pr = q % FP, rr = r0 + pr;
    if (rr >= g.nr) continue;
    const double r2 = rsq[rr];
    const double2 A = slot(0)[q], B = slot(1)[q], C = slot(2)[q], D = slot(3)[q], E = slot(4)[q], F = slot(5)[q];
    const double2 P = slot(6)[q], Q = slot(7)[q];
    double2 xr = make_double2(B.x * F.x - C.x * E.x, B.y * F.y - C.y * E.y);
    double2 xp = make_double2(C.x * D.x - A.x * F.x, C.y * D.y - A.y * F.y);
    double2 xz = make_double2((A.x * E.x - B.x * D.x) / r2, (A.y * E.y - B.y * D.y) / r2);
    double2 tt = make_double2(A.x * P.x + B.x * Q.x, A.y * P.y + B.y * Q.y);
    if (linear) {
      xr = make_double2(xr.x + two_omega * B.x, xr.y + two_omega * B.y);
      xp = make_double2(xp.x - two_omega * A.x, xp.y - two_omega * A.y);
    }
    slot(3)[q] = xr;
    slot(4)[q] = xp;
    slot(5)[q] = xz;
    slot(7)[q] = tt;
  }
  __syncthreads();
  // phase 2: forward X, Y (Z if no linear terms); then inverse bz -> slot 6
  // (and W -> slot 0, rur being no longer needed)
  {
    const bool fwd = grp < (linear ? 2 : 3);
    fz_forward<R1, R2>(fwd, g, slot(3 + (fwd ? grp : 0)), io.out[fwd ? grp : 0], r0, z, p, t, gt, tw, twh);
  }
  {
    const bool inv = (grp == 3) || (linear && grp == 4);
    fz_inverse<R1, R2>(inv, g, io.in[grp == 4 ? 9 : 8], slot(grp == 4 ? 0 : 6), r0, z, p, t, tw, twh);
  }
  // products B
  for (int q = threadIdx.x; q < SL; q += 256) {
    const int pr = q % FP, rr = r0 + pr;
    if (rr >= g.nr) continue;
    const double r2 = rsq[rr];
    const double2 C = slot(2)[q], R = slot(6)[q], T = slot(7)[q];
    const double vr = T.x / r2 + C.x * R.x, vi = T.y / r2 + C.y * R.y;
    double2 nb = make_double2(-vr, -vi);
    if (linear) {
      const double2 W = slot(0)[q], Z = slot(5)[q];
      nb = make_double2(nb.x + n2 * C.x, nb.y + n2 * C.y);
      slot(5)[q] = make_double2(Z.x - W.x, Z.y - W.y);
    }
    slot(7)[q] = nb;
  }
  __syncthreads();
  // phase 3: forward N_b (and N_z with linear terms)
  {
    const bool act = (grp == 0) || (linear && grp == 1);
    fz_forward<R1, R2>(act, g, slot(grp == 1 ? 5 : 7), io.out[grp == 1 ? 2 : 3], r0, z, p, t, gt, tw, twh);
  }
}

// ---------------------------------------------------------------------------
// NTH = 256: register-resident variant.
//
// 32 lanes own one theta line (two lines per wave, eight adjacent radial
// nodes per block).  Each lane keeps 8 of the line's 256 physical points in
// registers, at the same positions for every field, so the products are
// formed in registers and only the transposes of the transforms use LDS.
// The 256-point transform is 8 x 8 x 4:
//   j = 32 a + b (lane b holds a = 0..7)      -> 8-point DFT over a, twiddle
//   exchange to lane 4 c + f (holds e = 0..7)  -> 8-point DFT over e, twiddle
//   exchange to lane 4 c + gh (holds f = 0..3 for g = gh, gh + 4)
//                                               -> 4-point DFTs over f
// so lane 4 c + gh ends with the points k = c + 8 g + 64 h, h = 0..3.  The
// forward transform runs the same stages backwards with conjugate roots.
// Waves never wait for one another; the LDS of a line is touched only by
// its own 32 lanes.
// ---------------------------------------------------------------------------
namespace w256 {

constexpr int N = 256;
constexpr int LP = 288;  // LDS double2 per line (two exchange layouts, padded against bank conflicts)
__device__ __forceinline__ int e1(int c, int b) { return c * 36 + b; }
__device__ __forceinline__ int e2(int c, int f, int g) { return c * 36 + f * 9 + g; }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane twiddles, copied to LDS once per block:
//   t1[32 c + b] = W_N^{b c}, t2[4 g + f] = W_N^{8 f g}, th[k] = exp(-i pi k / N)
struct Ctx {
  const double2 *t1, *t2, *th;
  double2 *s;                       // LDS of this line
  int b;                            // lane within the line
};

// 8- and 4-point DFTs with the roots exp(S 2 pi i k / R) written out:
// multiplications by 1 and +-i are exact moves, exp(+-i pi/4) uses sqrt(1/2).
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
template <int S>
__device__ __forceinline__ double2 mul_si(double2 d) {  // (S i) d
  return S > 0 ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
}
template <int S>
__device__ __forceinline__ void dft4s(double2 (&y)[4]) {
  const double2 a = cadd(y[0], y[2]), b = csub(y[0], y[2]), c = cadd(y[1], y[3]), d = mul_si<S>(csub(y[1], y[3]));
  y[0] = cadd(a, c);
  y[1] = cadd(b, d);
  y[2] = csub(a, c);
  y[3] = csub(b, d);
}
template <int S>
__device__ __forceinline__ void dft8s(double2 (&x)[8]) {
  constexpr double h = 0.70710678118654752440;  // sqrt(1/2)
  double2 e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
  dft4s<S>(e);
  dft4s<S>(o);
  double2 t[4];
  t[0] = o[0];
  t[1] = S > 0 ? make_double2(h * (o[1].x - o[1].y), h * (o[1].y + o[1].x))
               : make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
  t[2] = mul_si<S>(o[2]);
  t[3] = S > 0 ? make_double2(h * (-o[3].x - o[3].y), h * (o[3].x - o[3].y))
               : make_double2(h * (o[3].y - o[3].x), h * (-o[3].y - o[3].x));
#pragma unroll
  for (int k = 0; k < 4; k++) {
    x[k] = cadd(e[k], t[k]);
    x[k + 4] = csub(e[k], t[k]);
  }
}

// dft8s for inputs with x[3] = x[4] = 0 (the unused slots of a spectrum
// with fewer than 96 modes): the additions of those zeros are left out.
template <int S>
__device__ __forceinline__ void dft8s_z34(double2 (&x)[8]) {
  constexpr double h = 0.70710678118654752440;
  // even part {x0, x2, 0, x6}, odd part {x1, 0, x5, x7}
  double2 e[4], o[4];
  {
    const double2 c = cadd(x[2], x[6]), d = mul_si<S>(csub(x[2], x[6]));
    e[0] = cadd(x[0], c);
    e[1] = cadd(x[0], d);
    e[2] = csub(x[0], c);
    e[3] = csub(x[0], d);
  }
  {
    const double2 a = cadd(x[1], x[5]), b = csub(x[1], x[5]), d = mul_si<-S>(x[7]);
    o[0] = cadd(a, x[7]);
    o[1] = cadd(b, d);
    o[2] = csub(a, x[7]);
    o[3] = csub(b, d);
  }
  double2 t[4];
  t[0] = o[0];
  t[1] = S > 0 ? make_double2(h * (o[1].x - o[1].y), h * (o[1].y + o[1].x))
               : make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
  t[2] = mul_si<S>(o[2]);
  t[3] = S > 0 ? make_double2(h * (-o[3].x - o[3].y), h * (o[3].x - o[3].y))
               : make_double2(h * (o[3].y - o[3].x), h * (-o[3].y - o[3].x));
#pragma unroll
  for (int k = 0; k < 4; k++) {
    x[k] = cadd(e[k], t[k]);
    x[k + 4] = csub(e[k], t[k]);
  }
}

__device__ __forceinline__ double2 cj(double2 w, bool inv) {
  if (inv) w.y = -w.y;
  return w;
}

// HORFFT backward of one PFP line (element m at src[lo + ndimr m], src the
// same for the whole block); lane b ends with the physical points
// P[c + 8 (gh + 4 s) + 64 h] in v[4 s + h], c = b >> 2, gh = b & 3.
// Needs NTCHOPDIM <= MMAX96: only a = 0..2 hold modes, a = 5..7 partners.
constexpr int MMAX96 = 96;
__device__ __forceinline__ void inverse(const Ctx &x, const KGrid &g, const double2 *__restrict__ src, int lo,
                                        bool live, double2 (&v)[8]) {
  const int b = x.b, mmax = g.ntchopdim;
  // branch-free: masked lanes load row 0 and discard it
  auto X = [&](int m, bool use) -> double2 {
    use = use && live && m < g.ntchop;
    double2 t = src[lo + (use ? g.ndimr * m : 0)];
    if (!use) t = make_double2(0.0, 0.0);
    if (m == 0) t.y = 0.0;
    return t;
  };
  // Hermitian merge: with e = X[j] + conj X[N-j], d = X[j] - conj X[N-j],
  // the point is e + i d conj(th[j]), th[j] = exp(-i pi j / N).  Only one of
  // the two terms is present: X[j] (j < 96) times 1 + i conj(th[j]), or
  // conj X[N-j] (j >= 160) times 1 - i conj(th[j]) = 1 + i th[N-j].
#pragma unroll
  for (int a = 0; a < 8; a++) {
    if (a == 3 || a == 4) {
      v[a] = make_double2(0.0, 0.0);
      continue;
    }
    const int j = 32 * a + b, jc = N - j;
    const double2 t = x.th[a < 3 ? j : jc];
    const double2 xv = a < 3 ? X(j, j < mmax) : fz_conj(X(jc, jc < mmax));
    const double2 q = a < 3 ? make_double2(1.0 + t.y, t.x) : make_double2(1.0 - t.y, t.x);
    v[a] = fz_cmul(xv, q);
  }
  dft8s_z34<1>(v);
  // twiddles W^0 = 1 multiply exactly
#pragma unroll
  for (int c = 1; c < 8; c++) v[c] = fz_cmul(v[c], cj(x.t1[32 * c + b], true));
  // exchange 1: lane b holds y_b[c]; lane 4 c + f takes y_{4 e + f}[c]
#pragma unroll
  for (int c = 0; c < 8; c++) x.s[e1(c, b)] = v[c];
  wave_sync();
  const int c0 = b >> 2, f0 = b & 3;
#pragma unroll
  for (int e = 0; e < 8; e++) v[e] = x.s[e1(c0, 4 * e + f0)];
  wave_sync();
  dft8s<1>(v);
#pragma unroll
  for (int gg = 1; gg < 8; gg++) v[gg] = fz_cmul(v[gg], cj(x.t2[4 * gg + f0], true));
  // exchange 2: lane 4 c + f holds z_{c,f}[g]; lane 4 c + gh takes f = 0..3 of g = gh, gh + 4
#pragma unroll
  for (int gg = 0; gg < 8; gg++) x.s[e2(c0, f0, gg)] = v[gg];
  wave_sync();
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int f = 0; f < 4; f++) t[sg][f] = x.s[e2(c0, f, f0 + 4 * sg)];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    dft4s<1>(t[sg]);
#pragma unroll
    for (int h = 0; h < 4; h++) v[4 * sg + h] = t[sg][h];
  }
}

// HORFFT forward of this lane's physical points v (positions as left by
// `inverse`) into one PFP line (element m at dst[lo + ndimr m]).
__device__ __forceinline__ void forward(const Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *__restrict__ dst,
                                        int lo, bool inr, bool live) {
  const int b = x.b, c0 = b >> 2, f0 = b & 3;
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
#pragma unroll
    for (int h = 0; h < 4; h++) t[sg][h] = v[4 * sg + h];
    dft4s<-1>(t[sg]);
    const int gg = f0 + 4 * sg;
#pragma unroll
    for (int f = 1; f < 4; f++) t[sg][f] = fz_cmul(t[sg][f], x.t2[4 * gg + f]);
  }
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int f = 0; f < 4; f++) x.s[e2(c0, f, f0 + 4 * sg)] = t[sg][f];
  wave_sync();
  double2 u[8];
#pragma unroll
  for (int gg = 0; gg < 8; gg++) u[gg] = x.s[e2(c0, f0, gg)];
  wave_sync();
  dft8s<-1>(u);
#pragma unroll
  for (int e = 0; e < 8; e++) x.s[e1(c0, 4 * e + f0)] = u[e];
  wave_sync();
#pragma unroll
  for (int c = 0; c < 8; c++) u[c] = x.s[e1(c, b)];
  wave_sync();
#pragma unroll
  for (int c = 1; c < 8; c++) u[c] = fz_cmul(u[c], x.t1[32 * c + b]);
  dft8s<-1>(u);
  // u[a] = Y[32 a + b]; the Hermitian split pairs Y[m] with Y[N - m],
  // which lane (32 - b) mod 32 holds at a' = 7 - a (lane 0: a' = (8 - a) mod 8)
  const double den = (double)(2 * N);
  const int src = (32 - b) & 31;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (32 * a >= g.ntchopdim) break;
    const int m = 32 * a + b;
    double2 pc;
    pc.x = __shfl(u[7 - a].x, src, 32);
    pc.y = __shfl(u[7 - a].y, src, 32);
    if (b == 0) pc = u[(8 - a) & 7];
    if (m < g.ntchopdim && inr) {
      double2 zm = u[a], zc = fz_conj(pc);
      double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
      double2 wd = fz_cmul(x.th[m], d);
      double2 o = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
      o = make_double2(o.x / den, o.y / den);
      if (m == 0) o.y = 0.0;
      if (m >= g.ntchop || !live) o = make_double2(0.0, 0.0);
      dst[lo + g.ndimr * m] = o;
    }
  }
}

}  // namespace w256

// Same products as k_fused_products, in the same floating-point operations.
// The outputs r N_r, r N_theta, N_z overwrite ror, rop, oz in the callers,
// so all three are read before the first of them is written.  The fields go
// to physical space one at a time; at most five stay in registers, and one
// finished product (Y) waits in LDS.
template <bool LINEAR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_fused_products_256(KGrid g, const double *__restrict__ rsq, const double2 *tw, const double2 *twh, FusedIO io,
                     double two_omega, double n2) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && io.zl) io.zl[0] = 0.0;
  using namespace w256;
  __shared__ double2 S[8 * LP];
  __shared__ double2 Y[8 * N];  // [line][point i][lane]
  __shared__ double2 T1[256], T2[32], TH[N + 1];
  const int nrb = (g.ndimr + 7) / 8;
  const int lid = xcd_logical(nrb * g.ndimx);
  if (lid < 0) return;
  {
    const int q = threadIdx.x;
    T1[q] = tw[(q & 31) * (q >> 5)];
    if (q < 32) T2[q] = tw[8 * (q & 3) * (q >> 2)];
    TH[q] = twh[q];
    if (q == 0) TH[N] = twh[N];
  }
  __syncthreads();
  const int z = lid / nrb, line = threadIdx.x >> 5;
  const int rr = (lid % nrb) * 8 + line;
  const bool inr = rr < g.ndimr, live = rr < g.nr;
  Ctx x;
  x.t1 = T1;
  x.t2 = T2;
  x.th = TH;
  x.s = S + line * LP;
  x.b = threadIdx.x & 31;
  double2 *ys = Y + line * N + x.b;
  const size_t off = (size_t)(rr - line) + (size_t)g.ndimr * g.ndimth * z;  // same for the block
  const int lo = inr ? line : 0;
  auto in = [&](int q) { return io.in[q] + off; };
  auto out = [&](int q) { return io.out[q] + off; };
  const double r2 = live ? rsq[rr] : 1.0;

  double2 A[8], B[8], C[8], U[8], V[8], T[8];
  inverse(x, g, in(0), lo, live, A);  // rur
  inverse(x, g, in(1), lo, live, B);  // rup
  inverse(x, g, in(2), lo, live, C);  // uz
  // oz: U = rup oz, V = rur oz
  inverse(x, g, in(5), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    U[i] = make_double2(B[i].x * T[i].x, B[i].y * T[i].y);
    V[i] = make_double2(A[i].x * T[i].x, A[i].y * T[i].y);
  }
  // ror: Y = uz ror - rur oz (to LDS); V = rup ror
  inverse(x, g, in(3), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    ys[32 * i] = make_double2(C[i].x * T[i].x - V[i].x, C[i].y * T[i].y - V[i].y);
    V[i] = make_double2(B[i].x * T[i].x, B[i].y * T[i].y);
  }
  // rop: X = rup oz - uz rop in U; Z = (rur rop - rup ror) / r^2 in V
  inverse(x, g, in(4), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    U[i] = make_double2(U[i].x - C[i].x * T[i].x, U[i].y - C[i].y * T[i].y);
    V[i] = make_double2((A[i].x * T[i].x - V[i].x) / r2, (A[i].y * T[i].y - V[i].y) / r2);
  }
  // ror, rop, oz are read: the outputs may now overwrite them
  if (LINEAR) {
#pragma unroll
    for (int i = 0; i < 8; i++) U[i] = make_double2(U[i].x + two_omega * B[i].x, U[i].y + two_omega * B[i].y);
  }
  forward(x, g, U, out(0), lo, inr, live);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    T[i] = ys[32 * i];
    if (LINEAR) T[i] = make_double2(T[i].x - two_omega * A[i].x, T[i].y - two_omega * A[i].y);
  }
  forward(x, g, T, out(1), lo, inr, live);
  if (LINEAR) {
    // W: N_z = Z - W
    inverse(x, g, in(9), lo, live, T);
#pragma unroll
    for (int i = 0; i < 8; i++) V[i] = make_double2(V[i].x - T[i].x, V[i].y - T[i].y);
  }
  forward(x, g, V, out(2), lo, inr, live);
  // rbr, rbp: A = rur rbr + rup rbp
  inverse(x, g, in(6), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) A[i] = make_double2(A[i].x * T[i].x, A[i].y * T[i].y);
  inverse(x, g, in(7), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) A[i] = make_double2(A[i].x + B[i].x * T[i].x, A[i].y + B[i].y * T[i].y);
  // bz: N_b = -(T / r^2 + uz bz) (+ Nbar^2 uz)
  inverse(x, g, in(8), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const double vr = A[i].x / r2 + C[i].x * T[i].x, vi = A[i].y / r2 + C[i].y * T[i].y;
    double2 nb = make_double2(-vr, -vi);
    if (LINEAR) nb = make_double2(nb.x + n2 * C[i].x, nb.y + n2 * C[i].y);
    A[i] = nb;
  }
  forward(x, g, A, out(3), lo, inr, live);
}

// ---------------------------------------------------------------------------
// NTH = 128: register-resident variant.
//
// 16 lanes own one theta line (four lines per wave, sixteen adjacent radial
// nodes per block).  Each lane keeps 8 of the line's 128 points of every
// field in registers.  The 128-point transform is 8 x 4 x 4:
//   j = 16 a + b (lane b holds a = 0..7)         -> 8-point DFT over a, twiddle
//   exchange: lane 4 c2 + f (b = 4 e + f) holds e = 0..3 of c = c2, c2 + 4
//                                                 -> 4-point DFTs over e, twiddle
//   exchange: lane 4 c2 + g holds f = 0..3 of c = c2, c2 + 4
//                                                 -> 4-point DFTs over f
// so lane 4 c2 + g ends with the points k = c + 8 g + 32 h, c = c2, c2 + 4,
// h = 0..3.  The forward transform runs the stages backwards with
// conjugate roots.  Same products, waves and LDS discipline as NTH = 256.
// ---------------------------------------------------------------------------
namespace w128 {

constexpr int N = 128;
constexpr int LP = 160;  // LDS double2 per line (the two exchange layouts, padded against bank conflicts)
constexpr int MMAX48 = 48;  // NTCHOPDIM bound: slots a = 3, 4 hold no mode
__device__ __forceinline__ int e1(int c, int b) { return c * 20 + b; }
__device__ __forceinline__ int e2(int c, int f, int g) { return c * 20 + f * 5 + g; }
using w256::cadd;
using w256::cj;
using w256::csub;
using w256::dft4s;
using w256::dft8s;
using w256::dft8s_z34;
using w256::wave_sync;
using w256::Ctx;  // t1[16 c + b] = W_N^{b c}, t2[4 g + f] = W_N^{8 f g}, th[k] = exp(-i pi k / N)

// HORFFT backward of one PFP line (element m at src[lo + ndimr m]); lane b
// ends with the physical points P[c + 8 g + 32 h] in v[4 s + h], c = c2 + 4 s,
// c2 = b >> 2, g = b & 3.
__device__ __forceinline__ void inverse(const Ctx &x, const KGrid &g, const double2 *__restrict__ src, int lo,
                                        bool live, double2 (&v)[8]) {
  const int b = x.b, mmax = g.ntchopdim;
  auto X = [&](int m, bool use) -> double2 {
    use = use && live && m < g.ntchop;
    double2 t = src[lo + (use ? g.ndimr * m : 0)];
    if (!use) t = make_double2(0.0, 0.0);
    if (m == 0) t.y = 0.0;
    return t;
  };
  // Hermitian merge with one term per point, as at NTH = 256
#pragma unroll
  for (int a = 0; a < 8; a++) {
    if (a == 3 || a == 4) {
      v[a] = make_double2(0.0, 0.0);
      continue;
    }
    const int j = 16 * a + b, jc = N - j;
    const double2 t = x.th[a < 3 ? j : jc];
    const double2 xv = a < 3 ? X(j, j < mmax) : fz_conj(X(jc, jc < mmax));
    const double2 q = a < 3 ? make_double2(1.0 + t.y, t.x) : make_double2(1.0 - t.y, t.x);
    v[a] = fz_cmul(xv, q);
  }
  dft8s_z34<1>(v);
#pragma unroll
  for (int c = 1; c < 8; c++) v[c] = fz_cmul(v[c], cj(x.t1[16 * c + b], true));
  // exchange 1: lane b holds y_b[c]; lane 4 c2 + f takes y_{4 e + f}[c], c = c2, c2 + 4
#pragma unroll
  for (int c = 0; c < 8; c++) x.s[e1(c, b)] = v[c];
  wave_sync();
  const int c2 = b >> 2, f0 = b & 3;
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int e = 0; e < 4; e++) t[sg][e] = x.s[e1(c2 + 4 * sg, 4 * e + f0)];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    dft4s<1>(t[sg]);
#pragma unroll
    for (int gg = 1; gg < 4; gg++) t[sg][gg] = fz_cmul(t[sg][gg], cj(x.t2[4 * gg + f0], true));
  }
  // exchange 2: lane 4 c2 + f holds z_{c,f}[g]; lane 4 c2 + g takes f = 0..3
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int gg = 0; gg < 4; gg++) x.s[e2(c2 + 4 * sg, f0, gg)] = t[sg][gg];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int f = 0; f < 4; f++) t[sg][f] = x.s[e2(c2 + 4 * sg, f, f0)];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    dft4s<1>(t[sg]);
#pragma unroll
    for (int h = 0; h < 4; h++) v[4 * sg + h] = t[sg][h];
  }
}

// HORFFT forward of this lane's physical points v (positions as left by
// `inverse`) into one PFP line (element m at dst[lo + ndimr m]).
__device__ __forceinline__ void forward(const Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *__restrict__ dst,
                                        int lo, bool inr, bool live) {
  const int b = x.b, c2 = b >> 2, g0 = b & 3;
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
#pragma unroll
    for (int h = 0; h < 4; h++) t[sg][h] = v[4 * sg + h];
    dft4s<-1>(t[sg]);
#pragma unroll
    for (int f = 1; f < 4; f++) t[sg][f] = fz_cmul(t[sg][f], x.t2[4 * g0 + f]);
  }
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int f = 0; f < 4; f++) x.s[e2(c2 + 4 * sg, f, g0)] = t[sg][f];
  wave_sync();
  // lane 4 c2 + f takes g = 0..3
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int gg = 0; gg < 4; gg++) t[sg][gg] = x.s[e2(c2 + 4 * sg, g0, gg)];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    dft4s<-1>(t[sg]);
#pragma unroll
    for (int e = 0; e < 4; e++) x.s[e1(c2 + 4 * sg, 4 * e + g0)] = t[sg][e];
  }
  wave_sync();
  double2 u[8];
#pragma unroll
  for (int c = 0; c < 8; c++) u[c] = x.s[e1(c, b)];
  wave_sync();
#pragma unroll
  for (int c = 1; c < 8; c++) u[c] = fz_cmul(u[c], x.t1[16 * c + b]);
  dft8s<-1>(u);
  // u[a] = Y[16 a + b]; Y[N - m] sits in lane (16 - b) mod 16 at a' = 7 - a
  // (lane 0: a' = (8 - a) mod 8)
  const double den = (double)(2 * N);
  const int src = (16 - b) & 15;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (16 * a >= g.ntchopdim) break;
    const int m = 16 * a + b;
    double2 pc;
    pc.x = __shfl(u[7 - a].x, src, 16);
    pc.y = __shfl(u[7 - a].y, src, 16);
    if (b == 0) pc = u[(8 - a) & 7];
    if (m < g.ntchopdim && inr) {
      double2 zm = u[a], zc = fz_conj(pc);
      double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
      double2 wd = fz_cmul(x.th[m], d);
      double2 o = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
      o = make_double2(o.x / den, o.y / den);
      if (m == 0) o.y = 0.0;
      if (m >= g.ntchop || !live) o = make_double2(0.0, 0.0);
      dst[lo + g.ndimr * m] = o;
    }
  }
}

}  // namespace w128

// ---------------------------------------------------------------------------
// NTH = 64: 8 lanes own one theta line (eight lines per wave), each keeping
// 8 of its 64 points.  The 64-point transform is 8 x 8 with one exchange:
//   j = 8 a + b (lane b holds a = 0..7) -> 8-point DFT over a, twiddle
//   exchange: lane c holds b = 0..7     -> 8-point DFT over b
// so lane c ends with the points k = c + 8 d, d = 0..7.
// ---------------------------------------------------------------------------
namespace w64 {

constexpr int N = 64;
constexpr int LP = 72;      // LDS double2 per line (rows of 9 against bank conflicts)
constexpr int MMAX24 = 24;  // NTCHOPDIM bound: slots a = 3, 4 hold no mode
__device__ __forceinline__ int e1(int c, int b) { return c * 9 + b; }
using w256::cj;
using w256::Ctx;  // t1[8 c + b] = W_N^{b c}, th[k] = exp(-i pi k / N)
using w256::dft8s;
using w256::dft8s_z34;
using w256::wave_sync;

__device__ __forceinline__ void inverse(const Ctx &x, const KGrid &g, const double2 *__restrict__ src, int lo,
                                        bool live, double2 (&v)[8]) {
  const int b = x.b, mmax = g.ntchopdim;
  auto X = [&](int m, bool use) -> double2 {
    use = use && live && m < g.ntchop;
    double2 t = src[lo + (use ? g.ndimr * m : 0)];
    if (!use) t = make_double2(0.0, 0.0);
    if (m == 0) t.y = 0.0;
    return t;
  };
#pragma unroll
  for (int a = 0; a < 8; a++) {
    if (a == 3 || a == 4) {
      v[a] = make_double2(0.0, 0.0);
      continue;
    }
    const int j = 8 * a + b, jc = N - j;
    const double2 t = x.th[a < 3 ? j : jc];
    const double2 xv = a < 3 ? X(j, j < mmax) : fz_conj(X(jc, jc < mmax));
    const double2 q = a < 3 ? make_double2(1.0 + t.y, t.x) : make_double2(1.0 - t.y, t.x);
    v[a] = fz_cmul(xv, q);
  }
  dft8s_z34<1>(v);
#pragma unroll
  for (int c = 1; c < 8; c++) v[c] = fz_cmul(v[c], cj(x.t1[8 * c + b], true));
#pragma unroll
  for (int c = 0; c < 8; c++) x.s[e1(c, b)] = v[c];
  wave_sync();
#pragma unroll
  for (int bb = 0; bb < 8; bb++) v[bb] = x.s[e1(b, bb)];
  wave_sync();
  dft8s<1>(v);
}

__device__ __forceinline__ void forward(const Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *__restrict__ dst,
                                        int lo, bool inr, bool live) {
  const int b = x.b;
  double2 u[8];
#pragma unroll
  for (int d = 0; d < 8; d++) u[d] = v[d];
  dft8s<-1>(u);
#pragma unroll
  for (int bb = 0; bb < 8; bb++) x.s[e1(b, bb)] = u[bb];
  wave_sync();
#pragma unroll
  for (int c = 0; c < 8; c++) u[c] = x.s[e1(c, b)];
  wave_sync();
#pragma unroll
  for (int c = 1; c < 8; c++) u[c] = fz_cmul(u[c], x.t1[8 * c + b]);
  dft8s<-1>(u);
  // u[a] = Y[8 a + b]; Y[N - m] sits in lane (8 - b) mod 8 at a' = 7 - a
  // (lane 0: a' = (8 - a) mod 8)
  const double den = (double)(2 * N);
  const int src = (8 - b) & 7;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (8 * a >= g.ntchopdim) break;
    const int m = 8 * a + b;
    double2 pc;
    pc.x = __shfl(u[7 - a].x, src, 8);
    pc.y = __shfl(u[7 - a].y, src, 8);
    if (b == 0) pc = u[(8 - a) & 7];
    if (m < g.ntchopdim && inr) {
      double2 zm = u[a], zc = fz_conj(pc);
      double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
      double2 wd = fz_cmul(x.th[m], d);
      double2 o = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
      o = make_double2(o.x / den, o.y / den);
      if (m == 0) o.y = 0.0;
      if (m >= g.ntchop || !live) o = make_double2(0.0, 0.0);
      dst[lo + g.ndimr * m] = o;
    }
  }
}

}  // namespace w64

// ---------------------------------------------------------------------------
// NTH = 32: 4 lanes own one theta line (sixteen lines per wave), each keeping
// 8 of its 32 points.  The 32-point transform is 8 x 4 with one exchange:
//   j = 4 a + b (lane b holds a = 0..7)          -> 8-point DFT over a, twiddle
//   exchange: lane L holds b = 0..3 of c = L, L + 4 -> 4-point DFTs over b
// so lane L ends with the points k = c + 8 d, c = L, L + 4, d = 0..3.
// ---------------------------------------------------------------------------
namespace w32 {

constexpr int N = 32;
constexpr int LP = 40;      // LDS double2 per line (rows of 5 against bank conflicts)
constexpr int MMAX12 = 12;  // NTCHOPDIM bound: slots a = 3, 4 hold no mode
__device__ __forceinline__ int e1(int c, int b) { return c * 5 + b; }
using w256::cj;
using w256::Ctx;  // t1[4 c + b] = W_N^{b c}, th[k] = exp(-i pi k / N)
using w256::dft4s;
using w256::dft8s;
using w256::dft8s_z34;
using w256::wave_sync;

__device__ __forceinline__ void inverse(const Ctx &x, const KGrid &g, const double2 *__restrict__ src, int lo,
                                        bool live, double2 (&v)[8]) {
  const int b = x.b, mmax = g.ntchopdim;
  auto X = [&](int m, bool use) -> double2 {
    use = use && live && m < g.ntchop;
    double2 t = src[lo + (use ? g.ndimr * m : 0)];
    if (!use) t = make_double2(0.0, 0.0);
    if (m == 0) t.y = 0.0;
    return t;
  };
#pragma unroll
  for (int a = 0; a < 8; a++) {
    if (a == 3 || a == 4) {
      v[a] = make_double2(0.0, 0.0);
      continue;
    }
    const int j = 4 * a + b, jc = N - j;
    const double2 t = x.th[a < 3 ? j : jc];
    const double2 xv = a < 3 ? X(j, j < mmax) : fz_conj(X(jc, jc < mmax));
    const double2 q = a < 3 ? make_double2(1.0 + t.y, t.x) : make_double2(1.0 - t.y, t.x);
    v[a] = fz_cmul(xv, q);
  }
  dft8s_z34<1>(v);
#pragma unroll
  for (int c = 1; c < 8; c++) v[c] = fz_cmul(v[c], cj(x.t1[4 * c + b], true));
#pragma unroll
  for (int c = 0; c < 8; c++) x.s[e1(c, b)] = v[c];
  wave_sync();
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++)
#pragma unroll
    for (int bb = 0; bb < 4; bb++) t[sg][bb] = x.s[e1(b + 4 * sg, bb)];
  wave_sync();
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
    dft4s<1>(t[sg]);
#pragma unroll
    for (int d = 0; d < 4; d++) v[4 * sg + d] = t[sg][d];
  }
}

__device__ __forceinline__ void forward(const Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *__restrict__ dst,
                                        int lo, bool inr, bool live) {
  const int b = x.b;
  double2 t[2][4];
#pragma unroll
  for (int sg = 0; sg < 2; sg++) {
#pragma unroll
    for (int d = 0; d < 4; d++) t[sg][d] = v[4 * sg + d];
    dft4s<-1>(t[sg]);
#pragma unroll
    for (int bb = 0; bb < 4; bb++) x.s[e1(b + 4 * sg, bb)] = t[sg][bb];
  }
  wave_sync();
  double2 u[8];
#pragma unroll
  for (int c = 0; c < 8; c++) u[c] = x.s[e1(c, b)];
  wave_sync();
#pragma unroll
  for (int c = 1; c < 8; c++) u[c] = fz_cmul(u[c], x.t1[4 * c + b]);
  dft8s<-1>(u);
  // u[a] = Y[4 a + b]; Y[N - m] sits in lane (4 - b) mod 4 at a' = 7 - a
  // (lane 0: a' = (8 - a) mod 8)
  const double den = (double)(2 * N);
  const int src = (4 - b) & 3;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (4 * a >= g.ntchopdim) break;
    const int m = 4 * a + b;
    double2 pc;
    pc.x = __shfl(u[7 - a].x, src, 4);
    pc.y = __shfl(u[7 - a].y, src, 4);
    if (b == 0) pc = u[(8 - a) & 7];
    if (m < g.ntchopdim && inr) {
      double2 zm = u[a], zc = fz_conj(pc);
      double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
      double2 wd = fz_cmul(x.th[m], d);
      double2 o = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
      o = make_double2(o.x / den, o.y / den);
      if (m == 0) o.y = 0.0;
      if (m >= g.ntchop || !live) o = make_double2(0.0, 0.0);
      dst[lo + g.ndimr * m] = o;
    }
  }
}

}  // namespace w32

// Line shapes of the register-resident kernel below: lanes per theta line,
// LDS per line, the NTCHOPDIM bound of the pruned first stage, and the
// per-lane twiddle tables (T1[LANES c + b] = W_N^{b c}; T2 of NTH = 128's
// middle stage, T2[4 g + f] = W_N^{8 f g}).
struct Line128 {
  static constexpr int N = 128, LANES = 16, LP = w128::LP, MMAX = w128::MMAX48, NT2 = 16;
  __device__ static void tables(int q, double2 *T1, double2 *T2, const double2 *tw) {
    if (q < 128) T1[q] = tw[(q & 15) * (q >> 4)];
    if (q < 16) T2[q] = tw[8 * (q & 3) * (q >> 2)];
  }
  __device__ static void inverse(const w256::Ctx &x, const KGrid &g, const double2 *src, int lo, bool live,
                                 double2 (&v)[8]) {
    w128::inverse(x, g, src, lo, live, v);
  }
  __device__ static void forward(const w256::Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *dst, int lo, bool inr,
                                 bool live) {
    w128::forward(x, g, v, dst, lo, inr, live);
  }
};
struct Line32 {
  static constexpr int N = 32, LANES = 4, LP = w32::LP, MMAX = w32::MMAX12, NT2 = 1;
  __device__ static void tables(int q, double2 *T1, double2 *, const double2 *tw) {
    if (q < 32) T1[q] = tw[(q & 3) * (q >> 2)];
  }
  __device__ static void inverse(const w256::Ctx &x, const KGrid &g, const double2 *src, int lo, bool live,
                                 double2 (&v)[8]) {
    w32::inverse(x, g, src, lo, live, v);
  }
  __device__ static void forward(const w256::Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *dst, int lo, bool inr,
                                 bool live) {
    w32::forward(x, g, v, dst, lo, inr, live);
  }
};
struct Line64 {
  static constexpr int N = 64, LANES = 8, LP = w64::LP, MMAX = w64::MMAX24, NT2 = 1;
  __device__ static void tables(int q, double2 *T1, double2 *, const double2 *tw) {
    if (q < 64) T1[q] = tw[(q & 7) * (q >> 3)];
  }
  __device__ static void inverse(const w256::Ctx &x, const KGrid &g, const double2 *src, int lo, bool live,
                                 double2 (&v)[8]) {
    w64::inverse(x, g, src, lo, live, v);
  }
  __device__ static void forward(const w256::Ctx &x, const KGrid &g, const double2 (&v)[8], double2 *dst, int lo, bool inr,
                                 bool live) {
    w64::forward(x, g, v, dst, lo, inr, live);
  }
};

// The NTH = 256 kernel's products and field order on the lines of NTH = 128
// (Line128: 16 lanes, 16 lines per block), NTH = 64 (Line64: 8 lanes, 32
// lines) or NTH = 32 (Line32: 4 lanes, 64 lines); a block takes adjacent
// radial nodes of one z plane.
template <class L, bool LINEAR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_fused_products_reg(KGrid g, const double *__restrict__ rsq, const double2 *tw, const double2 *twh, FusedIO io,
                     double two_omega, double n2) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && io.zl) io.zl[0] = 0.0;
  constexpr int N = L::N, LANES = L::LANES, LINES = 256 / LANES;
  __shared__ double2 S[LINES * L::LP];
  __shared__ double2 Y[LINES * N];  // [line][point i][lane]
  __shared__ double2 T1[N], T2[L::NT2], TH[N + 1];
  const int nrb = (g.ndimr + LINES - 1) / LINES;
  const int lid = xcd_logical(nrb * g.ndimx);
  if (lid < 0) return;
  {
    const int q = threadIdx.x;
    L::tables(q, T1, T2, tw);
    if (q <= N) TH[q] = twh[q];
  }
  __syncthreads();
  const int z = lid / nrb, line = threadIdx.x / LANES;
  const int rr = (lid % nrb) * LINES + line;
  const bool inr = rr < g.ndimr, live = rr < g.nr;
  w256::Ctx x;
  x.t1 = T1;
  x.t2 = T2;
  x.th = TH;
  x.s = S + line * L::LP;
  x.b = threadIdx.x % LANES;
  double2 *ys = Y + line * N + x.b;
  const size_t off = (size_t)(rr - line) + (size_t)g.ndimr * g.ndimth * z;  // same for the block
  const int lo = inr ? line : 0;
  auto in = [&](int q) { return io.in[q] + off; };
  auto out = [&](int q) { return io.out[q] + off; };
  const double r2 = live ? rsq[rr] : 1.0;

  double2 A[8], B[8], C[8], U[8], V[8], T[8];
  L::inverse(x, g, in(0), lo, live, A);  // rur
  L::inverse(x, g, in(1), lo, live, B);  // rup
  L::inverse(x, g, in(2), lo, live, C);  // uz
  // oz: U = rup oz, V = rur oz
  L::inverse(x, g, in(5), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    U[i] = make_double2(B[i].x * T[i].x, B[i].y * T[i].y);
    V[i] = make_double2(A[i].x * T[i].x, A[i].y * T[i].y);
  }
  // ror: Y = uz ror - rur oz (to LDS); V = rup ror
  L::inverse(x, g, in(3), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    ys[LANES * i] = make_double2(C[i].x * T[i].x - V[i].x, C[i].y * T[i].y - V[i].y);
    V[i] = make_double2(B[i].x * T[i].x, B[i].y * T[i].y);
  }
  // rop: X = rup oz - uz rop in U; Z = (rur rop - rup ror) / r^2 in V
  L::inverse(x, g, in(4), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    U[i] = make_double2(U[i].x - C[i].x * T[i].x, U[i].y - C[i].y * T[i].y);
    V[i] = make_double2((A[i].x * T[i].x - V[i].x) / r2, (A[i].y * T[i].y - V[i].y) / r2);
  }
  // ror, rop, oz are read: the outputs may now overwrite them
  if (LINEAR) {
#pragma unroll
    for (int i = 0; i < 8; i++) U[i] = make_double2(U[i].x + two_omega * B[i].x, U[i].y + two_omega * B[i].y);
  }
  L::forward(x, g, U, out(0), lo, inr, live);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    T[i] = ys[LANES * i];
    if (LINEAR) T[i] = make_double2(T[i].x - two_omega * A[i].x, T[i].y - two_omega * A[i].y);
  }
  L::forward(x, g, T, out(1), lo, inr, live);
  if (LINEAR) {
    // W: N_z = Z - W
    L::inverse(x, g, in(9), lo, live, T);
#pragma unroll
    for (int i = 0; i < 8; i++) V[i] = make_double2(V[i].x - T[i].x, V[i].y - T[i].y);
  }
  L::forward(x, g, V, out(2), lo, inr, live);
  // rbr, rbp: A = rur rbr + rup rbp
  L::inverse(x, g, in(6), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) A[i] = make_double2(A[i].x * T[i].x, A[i].y * T[i].y);
  L::inverse(x, g, in(7), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) A[i] = make_double2(A[i].x + B[i].x * T[i].x, A[i].y + B[i].y * T[i].y);
  // bz: N_b = -(T / r^2 + uz bz) (+ Nbar^2 uz)
  L::inverse(x, g, in(8), lo, live, T);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const double vr = A[i].x / r2 + C[i].x * T[i].x, vi = A[i].y / r2 + C[i].y * T[i].y;
    double2 nb = make_double2(-vr, -vi);
    if (LINEAR) nb = make_double2(nb.x + n2 * C[i].x, nb.y + n2 * C[i].y);
    A[i] = nb;
  }
  L::forward(x, g, A, out(3), lo, inr, live);
}

bool fused_products_supported(const mlg_ctx *c) {
  const int n = c->g.nth;
  return n >= 2 && n <= 256 && (n & (n - 1)) == 0;
}

int launch_fused_products(mlg_ctx *c, const double2 *const in[10], double2 *const out[4], double two_omega, double n2,
                          int linear, double *zl) {
  KGrid g = kgrid(c);
  FusedIO io;
  for (int i = 0; i < 10; i++) io.in[i] = in[i];
  for (int i = 0; i < 4; i++) io.out[i] = out[i];
  io.zl = zl;
  // PFP traffic: 9 (10) live input fields read, 4 written
  KScope ks(c, K_PRODUCTS, 16.0 * g.nr * g.ntchopdim * (double)g.nx * ((linear ? 10 : 9) + 4));
#define L_REG(L_)                                                                                                    \
  do {                                                                                                               \
    dim3 grid(xcd_grid((g.ndimr + 256 / L_::LANES - 1) / (256 / L_::LANES) * g.ndimx));                             \
    if (linear)                                                                                                      \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fused_products_reg<L_, true>), grid, dim3(256), 0, c->stream, g, c->d.r,  \
                         c->d.tw_t, c->d.tw_h, io, two_omega, n2);                                                   \
    else                                                                                                             \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fused_products_reg<L_, false>), grid, dim3(256), 0, c->stream, g, c->d.r, \
                         c->d.tw_t, c->d.tw_h, io, two_omega, n2);                                                   \
    MLG_CHECK_HIP(hipGetLastError());                                                                                \
    return 0;                                                                                                        \
  } while (0)
  if (g.nth == 128 && g.ntchopdim <= Line128::MMAX) L_REG(Line128);
  if (g.nth == 64 && g.ntchopdim <= Line64::MMAX) L_REG(Line64);
  if (g.nth == 32 && g.ntchopdim <= Line32::MMAX) L_REG(Line32);
#undef L_REG
  if (g.nth == 256 && g.ntchopdim <= w256::MMAX96) {
    dim3 grid(xcd_grid((g.ndimr + 7) / 8 * g.ndimx));
    if (linear)
      hipLaunchKernelGGL(k_fused_products_256<true>, grid, dim3(256), 0, c->stream, g, c->d.r, c->d.tw_t, c->d.tw_h, io,
                         two_omega, n2);
    else
      hipLaunchKernelGGL(k_fused_products_256<false>, grid, dim3(256), 0, c->stream, g, c->d.r, c->d.tw_t, c->d.tw_h,
                         io, two_omega, n2);
    MLG_CHECK_HIP(hipGetLastError());
    return 0;
  }
  dim3 grid(xcd_grid((g.ndimr + FP - 1) / FP * g.ndimx));
#define L_FP(a_, b_)                                                                                                \
  hipLaunchKernelGGL((k_fused_products<a_, b_>), grid, dim3(256), 0, c->stream, g, c->d.r, c->d.tw_t, c->d.tw_h, io, \
                     two_omega, n2, linear)
  switch (g.nth) {
    case 2: L_FP(2, 1); break;
    case 4: L_FP(4, 1); break;
    case 8: L_FP(8, 1); break;
    case 16: L_FP(16, 1); break;
    case 64: L_FP(8, 8); break;
    case 32: L_FP(16, 2); break;
    case 128: L_FP(16, 8); break;
    case 256: L_FP(16, 16); break;
    default: return MLG_E_ARG;
  }
#undef L_FP
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
