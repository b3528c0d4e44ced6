// Axial (z) and azimuthal (theta) FFTs with the pencil transposes fused in.
//
// VERFFT (src/mod_fft.f90:928-1113): complex FFT of length NX along z for
// every (r, m), forward scaled by 1/NX and chopped to k = 0..NXCHOP-1 and
// -(NXCHOP-1)..-1, backward unchopped and unscaled.  The reference's two
// all-to-all exchanges become the load/store pattern of one kernel: a block
// owns P consecutive radial nodes of one m, loads their z pencils with
// coalesced r-contiguous rows into LDS, transforms, and writes them back
// r-contiguous in the other layout.
//
// HORFFT (src/mod_fft.f90:771-926): real FFT of length 2*NTH along theta for
// every (r, z).  A PPP slot j holds the physical values at theta_r(j) (real
// part) and theta_i(j) (imaginary part), i.e. the interleaved real sequence
// c[2j], c[2j+1]; that is exactly the packing of a real length-2N transform
// into a complex length-N one, so the slots are transformed as they lie and
// the half-spectrum is split/merged with the e^{-i pi m/NTH} twiddles.
//
// Power-of-two lengths: four-step register kernels up to 256, radix-2 in
// LDS above.  Other lengths (the reference plans FFTW at any length,
// src/mod_fft.f90:833-920,1006-1091): direct DFT sums over pencils staged in
// LDS, the phase index k*z mod N stepped through a table e^{-2 pi i j/N}.
#include "mlegs_internal.h"
#include "xcd.h"

namespace mlg {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

__device__ __forceinline__ int bitrev(int v, int lg) { return (int)(__brev((unsigned)v) >> (32 - lg)); }

// In-place radix-2 DIT over P pencils of length N stored [p][N] in LDS,
// input already in bit-reversed order.  tw[j] = e^{-2 pi i j/N}; sign > 0
// uses the conjugate twiddles.
__device__ void lds_fft(double2 *s, int P, int N, int lg, const double2 *__restrict__ tw, int sign) {
  const int nb = P * (N >> 1);
  for (int st = 0; st < lg; st++) {
    const int half = 1 << st;
    const int tstep = N >> (st + 1);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      int p = b / (N >> 1), q = b % (N >> 1);
      int grp = q / half, pos = q - grp * half;
      int i0 = p * N + grp * 2 * half + pos, i1 = i0 + half;
      double2 w = tw[pos * tstep];
      if (sign > 0) w.y = -w.y;
      double2 u = s[i0], v = cmul(s[i1], w);
      s[i0] = make_double2(u.x + v.x, u.y + v.y);
      s[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

static int pencils_for(int n) {
  // 64 KiB of LDS per block
  int p = 4096 / n;
  if (p < 1) p = 1;
  if (p > 32) p = 32;
  return p;
}

// ---- VERFFT forward: PFP (r, m<NTCHOPDIM, z) -> PFF (r, m, k)
__global__ void __launch_bounds__(256) k_verfft_fwd(KGrid g, int P, int lg, const double2 *__restrict__ tw,
                                                    const double2 *__restrict__ in, double2 *__restrict__ out) {
  extern __shared__ double2 s[];
  const int mm = blockIdx.y, r0 = blockIdx.x * P, N = g.nx;
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    int r = idx % P, z = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < g.nr) v = in[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)];
    s[r * N + bitrev(z, lg)] = v;
  }
  __syncthreads();
  lds_fft(s, P, N, lg, tw, -1);
  const double inv = (double)N;
  for (int idx = threadIdx.x; idx < P * g.nxchopdim; idx += blockDim.x) {
    int r = idx % P, kk = idx / P;
    if (r0 + r >= g.ndimr) continue;
    int src = kk < g.nxchop ? kk : (g.nxchoph - 1 + (kk - g.nxchop));
    double2 v = s[r * N + src];
    out[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)] = make_double2(v.x / inv, v.y / inv);
  }
}

// ---- VERFFT backward: PFF (r, m, k) -> PFP (r, m<NTCHOPDIM, z)
__global__ void __launch_bounds__(256) k_verfft_bwd(KGrid g, int P, int lg, const double2 *__restrict__ tw,
                                                    const double2 *__restrict__ in, double2 *__restrict__ out) {
  extern __shared__ double2 s[];
  const int mm = blockIdx.y, r0 = blockIdx.x * P, N = g.nx;
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) s[idx] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * g.nxchopdim; idx += blockDim.x) {
    int r = idx % P, kk = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < g.nr) v = in[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)];
    int z = kk < g.nxchop ? kk : (g.nxchoph - 1 + (kk - g.nxchop));
    s[r * N + bitrev(z, lg)] = v;
  }
  __syncthreads();
  lds_fft(s, P, N, lg, tw, +1);
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    int r = idx % P, z = idx / P;
    if (r0 + r >= g.ndimr) continue;
    out[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = s[r * N + z];
  }
}

static int launch_verfft_fwd_r2(mlg_ctx *c, const double2 *in, double2 *out) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nx), lg = ilog2(g.nx);
  dim3 grid((g.ndimr + P - 1) / P, g.ntchopdim);
  size_t lds = sizeof(double2) * (size_t)P * g.nx;
  KScope ks(c, K_VERFFT_FWD, 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
  hipLaunchKernelGGL(k_verfft_fwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_z, in, out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

static int launch_verfft_bwd_r2(mlg_ctx *c, const double2 *in, double2 *out) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nx), lg = ilog2(g.nx);
  dim3 grid((g.ndimr + P - 1) / P, g.ntchopdim);
  size_t lds = sizeof(double2) * (size_t)P * g.nx;
  KScope ks(c, K_VERFFT_BWD, 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
  hipLaunchKernelGGL(k_verfft_bwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_z, in, out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---- HORFFT forward (in place): PPP slots -> PFP m < NTCHOPDIM
// Z = FFT_NTH(slots); X_m = (Z_m + conj Z_{NTH-m})/2 - (i/2) W^m (Z_m - conj Z_{NTH-m}),
// W = e^{-i pi/NTH}; scaled by 1/(2 NTH); m >= NTCHOP dropped; X_0 made real.
__global__ void __launch_bounds__(256) k_horfft_fwd(KGrid g, int P, int lg, const double2 *__restrict__ tw,
                                                    const double2 *__restrict__ twh, double2 *__restrict__ a) {
  extern __shared__ double2 s[];
  const int z = blockIdx.y, r0 = blockIdx.x * P, N = g.nth;
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    int r = idx % P, j = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < g.nr) v = a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z)];
    s[r * N + bitrev(j, lg)] = v;
  }
  __syncthreads();
  lds_fft(s, P, N, lg, tw, -1);
  const double den = (double)(2 * N);
  const int mmax = g.ntchopdim;
  for (int idx = threadIdx.x; idx < P * mmax; idx += blockDim.x) {
    int r = idx % P, mm = idx / P;
    if (r0 + r >= g.ndimr) continue;
    double2 zm = s[r * N + (mm % N)], zc = cconj(s[r * N + ((N - mm) % N)]);
    double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
    double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
    double2 wd = cmul(twh[mm], d);                 // W^m (Z_m - conj Z_{N-m})
    double2 x = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);  // e - (i/2) wd
    x = make_double2(x.x / den, x.y / den);
    if (mm == 0) x.y = 0.0;
    if (mm >= g.ntchop) x = make_double2(0.0, 0.0);
    if (r0 + r >= g.nr) x = make_double2(0.0, 0.0);
    a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = x;
  }
}

// ---- HORFFT backward (in place): PFP m < NTCHOPDIM -> PPP slots
// E_m = X_m + conj X_{NTH-m}, O_m = (X_m - conj X_{NTH-m}) e^{+i pi m/NTH},
// slots = IFFT_NTH(E + i O) (unscaled); X_0 real, X_m = 0 for m >= NTCHOPDIM.
__global__ void __launch_bounds__(256) k_horfft_bwd(KGrid g, int P, int lg, const double2 *__restrict__ tw,
                                                    const double2 *__restrict__ twh, double2 *__restrict__ a) {
  extern __shared__ double2 sm[];
  const int z = blockIdx.y, r0 = blockIdx.x * P, N = g.nth;
  double2 *X = sm;              // [P][mmax]
  double2 *s = sm + (size_t)P * g.ntchopdim;  // [P][N]
  const int mmax = g.ntchopdim;
  for (int idx = threadIdx.x; idx < P * mmax; idx += blockDim.x) {
    int r = idx % P, mm = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < g.nr && mm < g.ntchop) v = a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)];
    if (mm == 0) v.y = 0.0;
    X[r * mmax + mm] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    int r = idx % P, j = idx / P;
    double2 xm = (j < mmax) ? X[r * mmax + j] : make_double2(0.0, 0.0);
    int jc = N - j;  // 1..N ; X_N (Nyquist) is chopped
    double2 xc = (jc < mmax && jc > 0) ? cconj(X[r * mmax + jc]) : make_double2(0.0, 0.0);
    if (j == 0) xc = make_double2(0.0, 0.0);  // X_N = 0
    double2 e = make_double2(xm.x + xc.x, xm.y + xc.y);
    double2 d = make_double2(xm.x - xc.x, xm.y - xc.y);
    double2 o = cmul(d, cconj(twh[j]));        // e^{+i pi j/N}
    s[r * N + bitrev(j, lg)] = make_double2(e.x - o.y, e.y + o.x);  // e + i o
  }
  __syncthreads();
  lds_fft(s, P, N, lg, tw, +1);
  for (int idx = threadIdx.x; idx < P * (N + 1); idx += blockDim.x) {
    int r = idx % P, j = idx / P;
    if (r0 + r >= g.ndimr) continue;
    double2 v = (j < N && r0 + r < g.nr) ? s[r * N + j] : make_double2(0.0, 0.0);
    a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z)] = v;
  }
}

static int launch_horfft_fwd_r2(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nth), lg = ilog2(g.nth);
  dim3 grid((g.ndimr + P - 1) / P, g.ndimx);
  size_t lds = sizeof(double2) * (size_t)P * g.nth;
  KScope ks(c, K_HORFFT_FWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_fwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_t, c->d.tw_h, a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

static int launch_horfft_bwd_r2(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nth + g.ntchopdim), lg = ilog2(g.nth);
  dim3 grid((g.ndimr + P - 1) / P, g.ndimx);
  size_t lds = sizeof(double2) * (size_t)P * (g.nth + g.ntchopdim);
  KScope ks(c, K_HORFFT_BWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_bwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_t, c->d.tw_h, a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// =========================================================================
// Four-step register FFTs for N = R1*R2 <= 256 (R1, R2 powers of two <= 16).
//   X[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} W_N^{n2 k1} sum_n1 x[R2 n1 + n2] W_R1^{n1 k1}
// Pass 1: thread (pencil r, n2) holds x[R2 n1 + n2], n1 < R1, in registers
// and does the R1-point DFT; the products with W_N^{n2 k1} go to LDS.
// Pass 2: thread (r, k1) reads its R2 values back and does the R2-point DFT.
// Lanes run over consecutive pencils (radial nodes), so every global access
// of a wave is a run of consecutive r.  Block: 256 threads, P = 256/TPP
// pencils; LDS holds P*N values.
// =========================================================================
template <int R>
struct Log2 {
  static constexpr int v = R <= 1 ? 0 : 1 + Log2<R / 2>::v;
};
template <>
struct Log2<1> {
  static constexpr int v = 0;
};

__device__ __forceinline__ constexpr int cbitrev(int i, int lg) {
  int r = 0;
  for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
  return r;
}

// In-register R-point DFT (natural order in and out); wR[j] = W_R^j (j < R/2),
// already conjugated for the inverse.
template <int R>
__device__ __forceinline__ void dft_reg(double2 (&v)[R], const double2 (&wR)[R / 2 > 0 ? R / 2 : 1]) {
  constexpr int LG = Log2<R>::v;
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int j = cbitrev(i, LG);
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
        double2 u = v[g0 + p], x = cmul(v[g0 + p + half], w);
        v[g0 + p] = make_double2(u.x + x.x, u.y + x.y);
        v[g0 + p + half] = make_double2(u.x - x.x, u.y - x.y);
      }
    }
  }
}

template <int R>
__device__ __forceinline__ void load_wr(double2 (&wR)[R / 2 > 0 ? R / 2 : 1], const double2 *__restrict__ tw, int N, bool inv) {
#pragma unroll
  for (int j = 0; j < (R / 2 > 0 ? R / 2 : 1); j++) {
    double2 w = (R > 1) ? tw[j * (N / R)] : make_double2(1.0, 0.0);
    if (inv) w.y = -w.y;
    wR[j] = w;
  }
}

template <int R1, int R2>
struct Fft4 {
  static constexpr int N = R1 * R2;
  static constexpr int TPP = R1 > R2 ? R1 : R2;
  static constexpr int P = 256 / TPP;
  // VERFFT backward: 512 threads, 32 radial nodes per block at N = 256, so
  // its PFF reads and PFP writes are 512-byte runs (measured faster than
  // 256-byte runs; the forward direction measured the other way)
  static constexpr int PB = 512 / TPP;
};

// ---- general lengths: DFT sums from LDS
static bool pow2_len(int n) { return n >= 1 && (n & (n - 1)) == 0; }
static int dft_pencils(int n) {
  int p = 4096 / n;
  return p < 1 ? 1 : (p > 32 ? 32 : p);
}
// the wavenumber of chopped index kk (k = 0..NXCHOP-1, then -(NXCHOP-1)..-1 as N + k)
__device__ __forceinline__ int kval_of(const KGrid &g, int kk) { return kk < g.nxchop ? kk : g.nx - (g.nxchopdim - kk); }

// VERFFT forward: PFP (r, m, z) -> PFF (r, m, k), / NX
__global__ void __launch_bounds__(256) k_verfft_fwd_dft(KGrid g, int P, const double2 *__restrict__ tw,
                                                        const double2 *__restrict__ in, double2 *__restrict__ out) {
  extern __shared__ double2 s[];
  const int mm = blockIdx.y, r0 = blockIdx.x * P, N = g.nx;
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    const int r = idx % P, z = idx / P;
    s[r * N + z] = (r0 + r < g.nr) ? in[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)]
                                   : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * g.nxchopdim; idx += blockDim.x) {
    const int r = idx % P, kk = idx / P;
    if (r0 + r >= g.ndimr) continue;
    const int k = kval_of(g, kk);
    double2 acc = make_double2(0.0, 0.0);
    int ph = 0;
    for (int z = 0; z < N; z++) {
      acc = cadd(acc, cmul(s[r * N + z], tw[ph]));
      ph += k;
      if (ph >= N) ph -= N;
    }
    out[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)] = make_double2(acc.x / N, acc.y / N);
  }
}

// VERFFT backward: PFF (r, m, k) -> PFP (r, m, z), unscaled
__global__ void __launch_bounds__(256) k_verfft_bwd_dft(KGrid g, int P, const double2 *__restrict__ tw,
                                                        const double2 *__restrict__ in, double2 *__restrict__ out) {
  extern __shared__ double2 s[];
  const int mm = blockIdx.y, r0 = blockIdx.x * P, N = g.nx, K = g.nxchopdim;
  for (int idx = threadIdx.x; idx < P * K; idx += blockDim.x) {
    const int r = idx % P, kk = idx / P;
    s[r * K + kk] = (r0 + r < g.nr) ? in[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)]
                                    : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    const int r = idx % P, z = idx / P;
    if (r0 + r >= g.ndimr) continue;
    double2 acc = make_double2(0.0, 0.0);
    int ph = 0;  // k z mod N, k = kk for kk < NXCHOP
    for (int kk = 0; kk < K; kk++) {
      if (kk == g.nxchop) ph = (int)(((long)kval_of(g, kk) * z) % N);
      acc = cadd(acc, cmul(s[r * K + kk], cconj(tw[ph])));
      ph += z;
      if (ph >= N) ph -= N;
    }
    out[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = acc;
  }
}

// HORFFT forward, in place: the 2 NTH real samples of a PPP row (slot j: Re at
// 2j, Im at 2j+1) -> X_m = sum_t s_t e^{-2 pi i m t/(2 NTH)} / (2 NTH) for
// m < NTCHOPDIM (zero from NTCHOP), X_0 real; results times s0 (m = 0) or s1.
__global__ void __launch_bounds__(256) k_horfft_fwd_dft(KGrid g, int P, const double2 *__restrict__ tw2, double2 *a,
                                                        double s0, double s1) {
  extern __shared__ double sr[];
  const int z = blockIdx.y, r0 = blockIdx.x * P, N = g.nth, N2 = 2 * N;
  for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
    const int r = idx % P, j = idx / P;
    double2 v = (r0 + r < g.nr) ? a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z)]
                                : make_double2(0.0, 0.0);
    sr[r * N2 + 2 * j] = v.x;
    sr[r * N2 + 2 * j + 1] = v.y;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * g.ntchopdim; idx += blockDim.x) {
    const int r = idx % P, mm = idx / P;
    if (r0 + r >= g.ndimr) continue;
    double2 x = make_double2(0.0, 0.0);
    if (mm < g.ntchop && r0 + r < g.nr) {
      int ph = 0;
      for (int t = 0; t < N2; t++) {
        const double v = sr[r * N2 + t];
        x = make_double2(x.x + v * tw2[ph].x, x.y + v * tw2[ph].y);
        ph += mm;
        if (ph >= N2) ph -= N2;
      }
      const double sm = (mm == 0 ? s0 : s1) / N2;
      x = make_double2(x.x * sm, x.y * sm);
      if (mm == 0) x.y = 0.0;
    }
    a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = x;
  }
}

// HORFFT backward, in place: s_t = Re X_0 + 2 sum_{1 <= m < NTCHOP} Re(X_m
// e^{+2 pi i m t/(2 NTH)}), slot j = (s_2j, s_2j+1), slot NTH = 0; inputs times
// s0 (m = 0) or s1 and, with wr, times wr(min(r, NR-1-r)).
__global__ void __launch_bounds__(256) k_horfft_bwd_dft(KGrid g, int P, const double2 *__restrict__ tw2, double2 *a,
                                                        double s0, double s1, const double *wr) {
  extern __shared__ double2 sc[];
  const int z = blockIdx.y, r0 = blockIdx.x * P, N = g.nth, N2 = 2 * N, M = g.ntchopdim;
  for (int idx = threadIdx.x; idx < P * M; idx += blockDim.x) {
    const int r = idx % P, mm = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < g.nr && mm < g.ntchop) {
      v = a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)];
      double sm = mm == 0 ? s0 : s1;
      if (wr) sm *= wr[r0 + r < g.nr - 1 - (r0 + r) ? r0 + r : g.nr - 1 - (r0 + r)];
      v = make_double2(v.x * sm, v.y * sm);
    }
    if (mm == 0) v.y = 0.0;
    sc[r * M + mm] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < P * (N + 1); idx += blockDim.x) {
    const int r = idx % P, j = idx / P;
    if (r0 + r >= g.ndimr) continue;
    double2 out = make_double2(0.0, 0.0);
    if (j < N && r0 + r < g.nr) {
      double se = sc[r * M].x, so = sc[r * M].x;
      int pe = 0, po = 0;  // m (2j) and m (2j+1) mod 2N
      for (int mm = 1; mm < M; mm++) {
        pe += 2 * j;
        while (pe >= N2) pe -= N2;
        po += 2 * j + 1;
        while (po >= N2) po -= N2;
        const double2 c = sc[r * M + mm];
        se += 2.0 * (c.x * tw2[pe].x + c.y * tw2[pe].y);
        so += 2.0 * (c.x * tw2[po].x + c.y * tw2[po].y);
      }
      out = make_double2(se, so);
    }
    a[(size_t)(r0 + r) + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z)] = out;
  }
}

static int launch_verfft_fwd_dft(mlg_ctx *c, const double2 *in, double2 *out) {
  KGrid g = kgrid(c);
  const int P = dft_pencils(g.nx);
  KScope ks(c, K_VERFFT_FWD, 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
  hipLaunchKernelGGL(k_verfft_fwd_dft, dim3((g.ndimr + P - 1) / P, g.ntchopdim), dim3(256),
                     sizeof(double2) * (size_t)P * g.nx, c->stream, g, P, c->d.tw_z, in, out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
static int launch_verfft_bwd_dft(mlg_ctx *c, const double2 *in, double2 *out) {
  KGrid g = kgrid(c);
  const int P = dft_pencils(g.nx);
  KScope ks(c, K_VERFFT_BWD, 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
  hipLaunchKernelGGL(k_verfft_bwd_dft, dim3((g.ndimr + P - 1) / P, g.ntchopdim), dim3(256),
                     sizeof(double2) * (size_t)P * g.nxchopdim, c->stream, g, P, c->d.tw_z, in, out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
static int launch_horfft_fwd_dft(mlg_ctx *c, double2 *a, double s0, double s1) {
  KGrid g = kgrid(c);
  const int P = dft_pencils(2 * g.nth);
  KScope ks(c, K_HORFFT_FWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_fwd_dft, dim3((g.ndimr + P - 1) / P, g.ndimx), dim3(256), sizeof(double) * 2 * (size_t)P * g.nth,
                     c->stream, g, P, c->d.tw_h2, a, s0, s1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
static int launch_horfft_bwd_dft(mlg_ctx *c, double2 *a, double s0, double s1, const double *wr) {
  KGrid g = kgrid(c);
  const int P = dft_pencils(2 * g.nth);
  KScope ks(c, K_HORFFT_BWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_bwd_dft, dim3((g.ndimr + P - 1) / P, g.ndimx), dim3(256),
                     sizeof(double2) * (size_t)P * g.ntchopdim, c->stream, g, P, c->d.tw_h2, a, s0, s1, wr);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

#define RET_FFT(x)             \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

// PFP slots m < NTCHOPDIM of a(r, m, z), r < NR, times s0 (m = 0) or s1
// (m >= 1), and times wr(min(r, NR-1-r)) when wr is given (the scalings of
// the adjoint transforms on the general-length theta path)
__global__ void k_scale_pfp_m(KGrid g, double2 *a, double s0, double s1, const double *wr) {
  const long n = (long)g.nr * g.ntchopdim * g.ndimx;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    const int r = (int)(idx % g.nr);
    const int j = (int)((idx / g.nr) % g.ntchopdim);
    const int z = (int)(idx / ((long)g.nr * g.ntchopdim));
    double sm = j == 0 ? s0 : s1;
    if (wr) sm *= wr[r < g.nr - 1 - r ? r : g.nr - 1 - r];
    double2 &v = a[(size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z)];
    v = make_double2(v.x * sm, v.y * sm);
  }
}
static int launch_scale_pfp_m(mlg_ctx *c, double2 *a, double s0, double s1, const double *wr) {
  KGrid g = kgrid(c);
  const long n = (long)g.nr * g.ntchopdim * g.ndimx;
  long b = (n + 255) / 256;
  hipLaunchKernelGGL(k_scale_pfp_m, dim3((unsigned)(b < 4096 ? b : 4096)), dim3(256), 0, c->stream, g, a, s0, s1, wr);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// Pass 1 + twiddle + LDS store + pass 2.  `v` enters holding x[R2 n1 + t]
// (thread t < R2) and leaves holding X[t + R1 k2] (thread t < R1).
template <int R1, int R2, int P = Fft4<R1, R2>::P>
__device__ __forceinline__ void fft4_core(double2 (&v1)[R1], double2 (&v2)[R2], double2 *s, int r, int t,
                                          const double2 *__restrict__ tw, bool inv) {
  constexpr int N = R1 * R2;
  if (t < R2) {
    double2 w1[R1 / 2 > 0 ? R1 / 2 : 1];
    load_wr<R1>(w1, tw, N, inv);
    dft_reg<R1>(v1, w1);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      double2 w = tw[t * k1];
      if (inv) w.y = -w.y;
      s[(k1 * R2 + t) * P + r] = (k1 == 0 || t == 0) ? v1[k1] : cmul(v1[k1], w);
    }
  }
  __syncthreads();
  if (t < R1) {
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) v2[n2] = s[(t * R2 + n2) * P + r];
    double2 w2[R2 / 2 > 0 ? R2 / 2 : 1];
    load_wr<R2>(w2, tw, N, inv);
    dft_reg<R2>(v2, w2);
  }
}

// VERFFT forward: PFP (r, m, z) -> PFF (r, m, k chopped), / NX.
template <int R1, int R2>
__global__ void __launch_bounds__(256) k_verfft_fwd4(KGrid g, const double2 *__restrict__ tw, TransformJobs jobs, int nj) {
  // logical block (radial chunk, mode m, field of the batch), radial chunks
  // fastest and grouped per XCD, so one XCD writes whole radial lines
  constexpr int N = R1 * R2, P = Fft4<R1, R2>::P;
  const int nrb = (g.ndimr + P - 1) / P;
  const int lid = xcd_logical(nrb * g.ntchopdim * nj);
  if (lid < 0) return;
  const int bx = lid % nrb, rest = lid / nrb, mm = rest % g.ntchopdim, fj = rest / g.ntchopdim;
  const double2 *__restrict__ in = jobs.in[fj];
  double2 *__restrict__ out = jobs.out[fj];
  __shared__ double2 s[P * N];
  const int r = threadIdx.x % P, t = threadIdx.x / P;
  const int rr = bx * P + r;
  double2 v1[R1], v2[R2];
  if (t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++)
      v1[n1] = (rr < g.nr) ? in[(size_t)rr + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * (R2 * n1 + t))]
                           : make_double2(0.0, 0.0);
  }
  fft4_core<R1, R2>(v1, v2, s, r, t, tw, false);
  if (t < R1 && rr < g.ndimr) {
    const double inv = (double)N;
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) {
      const int k = t + R1 * k2;
      int kk = -1;
      if (k < g.nxchop) kk = k;
      else if (k > N - g.nxchop) kk = g.nxchopdim + (k - N);
      if (kk >= 0)
        out[(size_t)rr + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)] = make_double2(v2[k2].x / inv, v2[k2].y / inv);
    }
  }
}

// VERFFT backward: PFF (r, m, k chopped) -> PFP (r, m, z), unscaled.
template <int R1, int R2>
__global__ void __launch_bounds__(512) k_verfft_bwd4(KGrid g, const double2 *__restrict__ tw, TransformJobs jobs, int nj,
                                                      const double *__restrict__ ak) {
  // logical block (radial chunk, mode m, field of the batch), radial chunks
  // fastest and grouped per XCD, so one XCD writes whole radial lines
  constexpr int N = R1 * R2, P = Fft4<R1, R2>::PB;
  const int nrb = (g.ndimr + P - 1) / P;
  const int lid = xcd_logical(nrb * g.ntchopdim * nj);
  if (lid < 0) return;
  const int bx = lid % nrb, rest = lid / nrb, mm = rest % g.ntchopdim, fj = rest / g.ntchopdim;
  const double2 *__restrict__ in = jobs.in[fj];
  double2 *__restrict__ out = jobs.out[fj];
  const bool dz = (jobs.dz >> fj) & 1u, dm = (jobs.dm >> fj) & 1u;
  if (lid == 0 && (int)threadIdx.x < jobs.nzl) jobs.zl[threadIdx.x][0] = 0.0;
  __shared__ double2 s[P * N];
  const int r = threadIdx.x % P, t = threadIdx.x / P;
  const int rr = bx * P + r;
  double2 v1[R1], v2[R2];
  if (t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
      const int k = R2 * n1 + t;
      int kk = -1;
      if (k < g.nxchop) kk = k;
      else if (k > N - g.nxchop) kk = g.nxchopdim + (k - N);
      v1[n1] = (rr < g.nr && kk >= 0) ? in[(size_t)rr + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * kk)]
                                      : make_double2(0.0, 0.0);
    }
    if (dz) {  // d/dz: i k a (GRAD's third component), after all loads are issued
#pragma unroll
      for (int n1 = 0; n1 < R1; n1++) {
        const int k = R2 * n1 + t;
        int kk = -1;
        if (k < g.nxchop) kk = k;
        else if (k > N - g.nxchop) kk = g.nxchopdim + (k - N);
        const double kv = kk >= 0 ? ak[mm + (size_t)g.ntchopdim * kk] : 0.0;
        v1[n1] = make_double2(-kv * v1[n1].y, kv * v1[n1].x);
      }
    }
    if (dm) {  // d/dtheta: i m a (GRAD's second component)
      const double mv = (double)(mm * g.minc);
#pragma unroll
      for (int n1 = 0; n1 < R1; n1++) v1[n1] = make_double2(-mv * v1[n1].y, mv * v1[n1].x);
    }
  }
  fft4_core<R1, R2, P>(v1, v2, s, r, t, tw, true);
  if (t < R1 && rr < g.ndimr) {
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++)
      out[(size_t)rr + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * (t + R1 * k2))] = v2[k2];
  }
}

// HORFFT forward, in place: PPP slots -> PFP m < NTCHOPDIM (split of the
// packed real transform as in k_horfft_fwd).
template <int R1, int R2>
__global__ void __launch_bounds__(256) k_horfft_fwd4(KGrid g, const double2 *__restrict__ tw, const double2 *__restrict__ twh,
                                                     HorJobs jobs) {
  constexpr int N = R1 * R2, P = Fft4<R1, R2>::P;
  double2 *__restrict__ a = jobs.a[blockIdx.z];
  __shared__ double2 s[P * N];
  const int r = threadIdx.x % P, t = threadIdx.x / P;
  const int z = blockIdx.y, r0 = blockIdx.x * P, rr = r0 + r;
  double2 v1[R1], v2[R2];
  if (t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++)
      v1[n1] = (rr < g.nr) ? a[(size_t)rr + (size_t)g.ndimr * ((size_t)(R2 * n1 + t) + (size_t)g.ndimth * z)]
                           : make_double2(0.0, 0.0);
  }
  fft4_core<R1, R2>(v1, v2, s, r, t, tw, false);
  __syncthreads();
  if (t < R1) {
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) s[(t + R1 * k2) * P + r] = v2[k2];
  }
  __syncthreads();
  const double den = (double)(2 * N);
  const int mmax = g.ntchopdim;
  for (int idx = threadIdx.x; idx < P * mmax; idx += 256) {
    const int pr = idx % P, mm = idx / P;
    if (r0 + pr >= g.ndimr) continue;
    double2 zm = s[(mm % N) * P + pr], zc = cconj(s[((N - mm) % N) * P + pr]);
    double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
    double2 d = make_double2(zm.x - zc.x, zm.y - zc.y);
    double2 wd = cmul(twh[mm], d);
    double2 x = make_double2(e.x + 0.5 * wd.y, e.y - 0.5 * wd.x);
    x = make_double2(x.x / den, x.y / den);
    const double sm = mm == 0 ? jobs.s0 : jobs.s1;
    x = make_double2(x.x * sm, x.y * sm);
    if (mm == 0) x.y = 0.0;
    if (mm >= g.ntchop || r0 + pr >= g.nr) x = make_double2(0.0, 0.0);
    a[(size_t)(r0 + pr) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)] = x;
  }
}

// HORFFT backward, in place: PFP m < NTCHOPDIM -> PPP slots (merge as in
// k_horfft_bwd, then an inverse length-NTH transform).
template <int R1, int R2>
__global__ void __launch_bounds__(256) k_horfft_bwd4(KGrid g, const double2 *__restrict__ tw, const double2 *__restrict__ twh,
                                                     HorJobs jobs) {
  constexpr int N = R1 * R2, P = Fft4<R1, R2>::P;
  double2 *__restrict__ a = jobs.a[blockIdx.z];
  __shared__ double2 s[P * N];
  const int r = threadIdx.x % P, t = threadIdx.x / P;
  const int z = blockIdx.y, r0 = blockIdx.x * P, rr = r0 + r;
  const int mmax = g.ntchopdim;
  for (int idx = threadIdx.x; idx < P * mmax; idx += 256) {
    const int pr = idx % P, mm = idx / P;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + pr < g.nr && mm < g.ntchop) {
      v = a[(size_t)(r0 + pr) + (size_t)g.ndimr * ((size_t)mm + (size_t)g.ndimth * z)];
      double sm = mm == 0 ? jobs.s0 : jobs.s1;
      if (jobs.wr) {
        const int rr = r0 + pr, rf = rr < g.nr - 1 - rr ? rr : g.nr - 1 - rr;
        sm = sm * jobs.wr[rf];
      }
      v = make_double2(v.x * sm, v.y * sm);
    }
    if (mm == 0) v.y = 0.0;
    s[mm * P + pr] = v;
  }
  __syncthreads();
  double2 v1[R1], v2[R2];
  if (t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) {
      const int j = R2 * n1 + t;
      double2 xm = (j < mmax) ? s[j * P + r] : make_double2(0.0, 0.0);
      const int jc = N - j;
      double2 xc = (jc < mmax && jc > 0 && j > 0) ? cconj(s[jc * P + r]) : make_double2(0.0, 0.0);
      double2 e = make_double2(xm.x + xc.x, xm.y + xc.y);
      double2 d = make_double2(xm.x - xc.x, xm.y - xc.y);
      double2 o = cmul(d, cconj(twh[j]));
      v1[n1] = make_double2(e.x - o.y, e.y + o.x);
    }
  }
  __syncthreads();
  fft4_core<R1, R2>(v1, v2, s, r, t, tw, true);
  if (rr < g.ndimr) {
    const bool live = rr < g.nr;
    if (t < R1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; k2++)
        a[(size_t)rr + (size_t)g.ndimr * ((size_t)(t + R1 * k2) + (size_t)g.ndimth * z)] = live ? v2[k2] : make_double2(0.0, 0.0);
    }
    if (t == 0) a[(size_t)rr + (size_t)g.ndimr * ((size_t)N + (size_t)g.ndimth * z)] = make_double2(0.0, 0.0);
  }
}

#define FFT4_SWITCH(N, CALL)                \
  switch (N) {                              \
    case 2: CALL(2, 1); break;              \
    case 4: CALL(4, 1); break;              \
    case 8: CALL(8, 1); break;              \
    case 16: CALL(16, 1); break;            \
    case 32: CALL(16, 2); break;            \
    case 64: CALL(8, 8); break;             \
    case 128: CALL(16, 8); break;           \
    case 256: CALL(16, 16); break;          \
    default: break;                         \
  }

int launch_verfft_fwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out) {
  if (n < 1 || n > MAX_TRANSFORM_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (g.nx > 256 || !pow2_len(g.nx)) {
    for (int i = 0; i < n; i++) {
      const int rc = pow2_len(g.nx) ? launch_verfft_fwd_r2(c, in[i], out[i]) : launch_verfft_fwd_dft(c, in[i], out[i]);
      if (rc) return rc;
    }
    return 0;
  }
  TransformJobs jobs;
  for (int i = 0; i < n; i++) {
    jobs.in[i] = in[i];
    jobs.ln[i] = nullptr;
    jobs.out[i] = out[i];
  }
  KScope ks(c, K_VERFFT_FWD, n * 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
#define L_VF(a, b)                                                                                                  \
  hipLaunchKernelGGL((k_verfft_fwd4<a, b>), dim3(xcd_grid((g.ndimr + Fft4<a, b>::P - 1) / Fft4<a, b>::P * g.ntchopdim * n)), \
                     dim3(256), 0, c->stream, g, c->d.tw_z, jobs, n)
  FFT4_SWITCH(g.nx, L_VF)
#undef L_VF
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_verfft_fwd(mlg_ctx *c, const double2 *in, double2 *out) { return launch_verfft_fwd_n(c, 1, &in, &out); }

int launch_verfft_bwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out, unsigned dz, unsigned dm,
                        double *const *zl, int nzl) {
  if (n < 1 || n > MAX_TRANSFORM_JOBS || nzl < 0 || nzl > MAX_TRANSFORM_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (g.nx > 256 || !pow2_len(g.nx)) {
    if (dz || dm) return MLG_E_ARG;  // derivatives only in the four-step kernel
    if (nzl > 0) {
      const int rc = launch_zero_lns(c, zl, nzl);
      if (rc) return rc;
    }
    for (int i = 0; i < n; i++) {
      const int rc = pow2_len(g.nx) ? launch_verfft_bwd_r2(c, in[i], out[i]) : launch_verfft_bwd_dft(c, in[i], out[i]);
      if (rc) return rc;
    }
    return 0;
  }
  TransformJobs jobs;
  for (int i = 0; i < n; i++) {
    jobs.in[i] = in[i];
    jobs.ln[i] = nullptr;
    jobs.out[i] = out[i];
  }
  jobs.dz = dz;
  jobs.dm = dm;
  jobs.nzl = nzl;
  for (int i = 0; i < nzl; i++) jobs.zl[i] = zl[i];
  KScope ks(c, K_VERFFT_BWD, n * 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
#define L_VB(a, b)                                                                                                    \
  hipLaunchKernelGGL((k_verfft_bwd4<a, b>), dim3(xcd_grid((g.ndimr + Fft4<a, b>::PB - 1) / Fft4<a, b>::PB * g.ntchopdim * n)), \
                     dim3(512), 0, c->stream, g, c->d.tw_z, jobs, n, c->d.ak)
  FFT4_SWITCH(g.nx, L_VB)
#undef L_VB
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_verfft_bwd(mlg_ctx *c, const double2 *in, double2 *out) { return launch_verfft_bwd_n(c, 1, &in, &out); }

int launch_horfft_fwd_n(mlg_ctx *c, int n, double2 *const *a, double s0, double s1) {
  if (n < 1 || n > MAX_HOR_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (!pow2_len(g.nth)) {
    for (int i = 0; i < n; i++) RET_FFT(launch_horfft_fwd_dft(c, a[i], s0, s1));
    return 0;
  }
  if (g.nth > 256) {
    for (int i = 0; i < n; i++) {
      RET_FFT(launch_horfft_fwd_r2(c, a[i]));
      if (s0 != 1.0 || s1 != 1.0) RET_FFT(launch_scale_pfp_m(c, a[i], s0, s1, nullptr));
    }
    return 0;
  }
  HorJobs jobs;
  for (int i = 0; i < n; i++) jobs.a[i] = a[i];
  jobs.s0 = s0, jobs.s1 = s1, jobs.wr = nullptr;
  KScope ks(c, K_HORFFT_FWD, n * 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
#define L_HF(a_, b_)                                                                                            \
  hipLaunchKernelGGL((k_horfft_fwd4<a_, b_>), dim3((g.ndimr + Fft4<a_, b_>::P - 1) / Fft4<a_, b_>::P, g.ndimx, n), \
                     dim3(256), 0, c->stream, g, c->d.tw_t, c->d.tw_h, jobs)
  FFT4_SWITCH(g.nth, L_HF)
#undef L_HF
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_horfft_fwd(mlg_ctx *c, double2 *a) { return launch_horfft_fwd_n(c, 1, &a, 1.0, 1.0); }

int launch_horfft_bwd_n(mlg_ctx *c, int n, double2 *const *a, double s0, double s1, const double *wr) {
  if (n < 1 || n > MAX_HOR_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (!pow2_len(g.nth)) {
    for (int i = 0; i < n; i++) RET_FFT(launch_horfft_bwd_dft(c, a[i], s0, s1, wr));
    return 0;
  }
  if (g.nth > 256) {
    for (int i = 0; i < n; i++) {
      if (s0 != 1.0 || s1 != 1.0 || wr) RET_FFT(launch_scale_pfp_m(c, a[i], s0, s1, wr));
      RET_FFT(launch_horfft_bwd_r2(c, a[i]));
    }
    return 0;
  }
  HorJobs jobs;
  for (int i = 0; i < n; i++) jobs.a[i] = a[i];
  jobs.s0 = s0, jobs.s1 = s1, jobs.wr = wr;
  KScope ks(c, K_HORFFT_BWD, n * 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
#define L_HB(a_, b_)                                                                                            \
  hipLaunchKernelGGL((k_horfft_bwd4<a_, b_>), dim3((g.ndimr + Fft4<a_, b_>::P - 1) / Fft4<a_, b_>::P, g.ndimx, n), \
                     dim3(256), 0, c->stream, g, c->d.tw_t, c->d.tw_h, jobs)
  FFT4_SWITCH(g.nth, L_HB)
#undef L_HB
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_horfft_bwd(mlg_ctx *c, double2 *a) { return launch_horfft_bwd_n(c, 1, &a, 1.0, 1.0, nullptr); }

}  // namespace mlg
