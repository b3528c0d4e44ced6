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
// Power-of-two lengths only (all configurations); radix-2 in LDS.
#include "mlegs_internal.h"

namespace mlg {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

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

int launch_verfft_fwd(mlg_ctx *c, const double2 *in, double2 *out) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nx), lg = ilog2(g.nx);
  dim3 grid((g.ndimr + P - 1) / P, g.ntchopdim);
  size_t lds = sizeof(double2) * (size_t)P * g.nx;
  KScope ks(c, K_VERFFT_FWD, 16.0 * g.nr * g.ntchopdim * (double)(g.nx + g.nxchopdim));
  hipLaunchKernelGGL(k_verfft_fwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_z, in, out);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_verfft_bwd(mlg_ctx *c, const double2 *in, double2 *out) {
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

int launch_horfft_fwd(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nth), lg = ilog2(g.nth);
  dim3 grid((g.ndimr + P - 1) / P, g.ndimx);
  size_t lds = sizeof(double2) * (size_t)P * g.nth;
  KScope ks(c, K_HORFFT_FWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_fwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_t, c->d.tw_h, a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_horfft_bwd(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  int P = pencils_for(g.nth + g.ntchopdim), lg = ilog2(g.nth);
  dim3 grid((g.ndimr + P - 1) / P, g.ndimx);
  size_t lds = sizeof(double2) * (size_t)P * (g.nth + g.ntchopdim);
  KScope ks(c, K_HORFFT_BWD, 16.0 * g.nr * g.nx * (double)(g.nth + g.ntchopdim));
  hipLaunchKernelGGL(k_horfft_bwd, grid, dim3(256), lds, c->stream, g, P, lg, c->d.tw_t, c->d.tw_h, a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
