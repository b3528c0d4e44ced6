// Mapped-Legendre radial transforms on FP64 MFMA (gfx950).
//
// RTRAN (src/mod_fft.f90:1150-1340) and the projection products of PROJECT
// (src/mod_legops.f90:2476-2529, EOMUL/OEMUL :2548-2648) are, for every
// azimuthal mode m, real x complex matrix products against the parity-folded
// table Pbar_n^m(x_i), i < NR/2.  All modes run in one grouped launch
// (every (m, tile) pair is one block, grouped per XCD); each mode's K
// extent follows the triangular truncation NRCHOPS(m).  Complex data enter as interleaved (re, im) real columns, so
// one real MFMA product serves both parts.
//
// MFMA: v_mfma_f64_16x16x4f64 -- lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; the 16x16 result holds C[l/16 + 4v][l%16] in register v
// (measured on gfx950; the f32 16x16x4 variant uses C[4(l/16)+v][l%16]).
#include <algorithm>

#include "mlegs_internal.h"
#include "xcd.h"

namespace mlg {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// LDS tiles are [row][col] with the column index XOR-swizzled by the row and
// a parity bit: column bits 0..3 ^= row bits 0..3, column bit 4 ^= row bit 0
// ^ parity, column bit 3 ^= parity.  The MFMA operand reads (16 consecutive
// columns of one row per 16 lanes) and the staging writes (16 rows, or 8
// rows of both parities, of one column per 16 lanes) then hit 16 different
// 8-byte bank pairs of a 32-bank LDS.
__device__ __forceinline__ int lsw(int r, int par) { return (r & 15) ^ (((r & 1) ^ par) << 4) ^ (par << 3); }

__device__ __forceinline__ int d_nrchops(int nrchop, int m) {
  int v = nrchop - m < nrchop ? nrchop - m : nrchop;
  return v > 0 ? v : 0;
}

// -------------------------------------------------------------------------
// Split hi/lo FP32 products (north-star precision mode; no counterpart in
// the reference, whose products are FP64 DGEMMs, src/mod_eig.f90:502-656).
// Within one K chunk a table row (scaled by 2^-E_row, stored split) and a
// data column (scaled by 2^-E_col when staged) are each a hi part on the
// 2^-10 grid plus an FP32 lo part, so both carry 34 significant bits
// relative to their row / column maximum over the chunk.  Per 4-deep MFMA
// step:
//   hh += A_hi B_hi                     (exact: 16 products of 21-bit
//                                        integers times one power of two sum
//                                        below 2^24 within a 16-deep chunk)
//   xx += A_hi B_lo + A_lo B_hi + A_lo B_lo   (FP32, ~2^-10 of the result)
// and after each K chunk the two FP32 accumulators are converted and added
// to the FP64 accumulator with the chunk's scale 2^(E_row + E_col).  The
// FP32 MFMA's output layout differs from the FP64 one: register v of lane l
// holds C[4(l/16) + v][l%16], four consecutive rows.
// -------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
// K chunk of every split product (pairs for RTRAN backward and the composed
// PROJECT, nodes for RTRAN forward and PROJECT): 16, so the hi x hi sums of
// a chunk stay exact
#define SPLIT_KC 16

template <bool SPLIT> struct LegElem { typedef double T; };
template <> struct LegElem<true> { typedef float2 T; };

__device__ __forceinline__ v4f mfma16x16x4f32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void mfma_split(float2 a, float2 b, v4f &hh, v4f &xx) {
  hh = mfma16x16x4f32(a.x, b.x, hh);
  xx = mfma16x16x4f32(a.x, b.y, xx);
  xx = mfma16x16x4f32(a.y, b.x, xx);
  xx = mfma16x16x4f32(a.y, b.y, xx);
}
// 2^e for -1022 <= e <= 1023
__device__ __forceinline__ double pow2i(int e) { return __longlong_as_double((long long)(1023 + e) << 52); }
// 2^e, or 0 below the normal range (a product of two scales of zero groups)
__device__ __forceinline__ double pow2c(int e) { return e < -1022 ? 0.0 : pow2i(min(e, 1023)); }
// biased exponent field of |x| (0 for zero)
__device__ __forceinline__ int bexp(double x) { return (int)((__double_as_longlong(x) >> 52) & 0x7ff); }
// the grid exponent E (|x| < 2^E for every x of the group) from the largest
// biased exponent of the group
__device__ __forceinline__ int grid_e(int emax) { return emax - 1022; }
// 2^-E, the factor that scales a group with grid exponent E into (-1, 1)
__device__ __forceinline__ double inv_grid(int e) { return pow2i(max(-e, -1022)); }
// x 2^-E split into hi (a multiple of 2^-10, |hi| <= 1) and lo = fl32(rest)
__device__ __forceinline__ float2 split_hl(double x, double inv_scale) {
  const double xn = x * inv_scale;
  const double hi = rint(xn * 1024.0) * (1.0 / 1024.0);
  return make_float2((float)hi, (float)(xn - hi));
}
// largest value of e over the lanes of an aligned group of `width` lanes
template <int width>
__device__ __forceinline__ int group_max(int e) {
#pragma unroll
  for (int o = 1; o < width; o <<= 1) e = max(e, __shfl_xor(e, o, 64));
  return e;
}
// add a chunk's FP32 accumulators into the FP64 one: register v holds the
// row with grid exponent er[v], the lane's column has grid exponent ec
__device__ __forceinline__ void fold_split(v4d &acc, v4f &hh, v4f &xx, int4 er, int ec) {
  const int e[4] = {er.x, er.y, er.z, er.w};
#pragma unroll
  for (int v = 0; v < 4; v++) acc[v] += ((double)hh[v] + (double)xx[v]) * pow2c(e[v] + ec);
  hh = (v4f){0.f, 0.f, 0.f, 0.f};
  xx = (v4f){0.f, 0.f, 0.f, 0.f};
}
// row of the 16x16 result held in register v of lane l
template <bool SPLIT>
__device__ __forceinline__ int acc_row(int lane, int v) {
  return SPLIT ? 4 * (lane >> 4) + v : (lane >> 4) + 4 * v;
}

// Split of an FP64 table into scaled (hi, lo) rows, one grid per K chunk.
// Row r holds the elements src[base(r) + k kstride], k < K, base(r) =
// (r / rin) ostride + (r % rin) istride; dst uses the same positions;
// rowe[r nch + c] = E of row r over chunk c (k in [c KC, (c+1) KC)).
__global__ void __launch_bounds__(256) k_split_rows(const double *__restrict__ src, float2 *__restrict__ dst,
                                                    int *__restrict__ rowe, long nrows, int rin, long ostride,
                                                    long istride, int K, long kstride, int KC) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  const long base = (r / rin) * ostride + (r % rin) * istride;
  const int nch = (K + KC - 1) / KC;
  for (int ch = 0; ch < nch; ch++) {
    const int k1 = min(K, (ch + 1) * KC);
    int emax = 0;
    for (int k = ch * KC; k < k1; k++) emax = max(emax, bexp(src[base + k * kstride]));
    const int e = grid_e(emax);
    const double inv = inv_grid(e);
    for (int k = ch * KC; k < k1; k++) dst[base + k * kstride] = split_hl(src[base + k * kstride], inv);
    rowe[r * nch + ch] = e;
  }
}

// Split mode: the lowest modes keep FP64 products (c->split_fp64_modes,
// at least m = 0 and 1).  PROJECT's V and D tables divide Pbar_n^m by
// (1 - x^2): singular at the poles for m = 0, 1, bounded for m = 2, so the
// low modes' products are large and cancel in psi and chi (the far field of
// a vortex, r u_theta -> const, against the log-term corrections at m = 0; a
// uniform drift at m = 1).  Split errors relative to those products measured
// 1e-8 relative to chi with FP64 at m = 0 only, 2e-9 with m < 2, and fall by
// ~4x per further mode.  g0 covers the FP64 modes, g1 the split ones.
static void split_modes(const mlg_ctx *c, const KGrid &g, KGrid &g0, KGrid &g1) {
  g0 = g;
  g1 = g;
  g0.mlo = 0;
  g0.mcnt = g.ntchopdim < c->split_fp64_modes ? g.ntchopdim : c->split_fp64_modes;
  g1.mlo = g0.mcnt;
  g1.mcnt = g.ntchopdim - g0.mcnt;
}
// sum of NRCHOPS(m) (p = 1) or NRCHOPS(m)^2 (p = 2) over the FP64 modes of split mode
static double split_fp64_sum(const KGrid &g0, int p) {
  double s = 0.0;
  for (int m = 0; m < g0.mcnt; m++) {
    const double n = nrchops_of(g0.nrchop, m * g0.minc);
    s += p == 1 ? n : n * n;
  }
  return s;
}

static int split_rows(mlg_ctx *c, const double *src, long n, float2 **dst, int **rowe, long nrows, int rin, long ostride,
                      long istride, int K, long kstride, int KC) {
  const long nch = (K + KC - 1) / KC;
  if (!*dst) MLG_CHECK_HIP(hipMalloc((void **)dst, sizeof(float2) * n));
  if (!*rowe) MLG_CHECK_HIP(hipMalloc((void **)rowe, sizeof(int) * nrows * nch));
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, c->stream, src, *dst, *rowe, nrows,
                     rin, ostride, istride, K, kstride, KC);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int build_split_tables(mlg_ctx *c) {
  KGrid g = kgrid(c);
  const long npf = (long)g.ntchopdim * g.npf * g.nrh, nvdt = 3L * g.ntchopdim * g.nrchopdim * g.nrh;
  int rc;
  // PF by node: row (m, i) runs over n with stride NRH; chunks of 32 n (16 pairs)
  if ((rc = split_rows(c, c->d.pf, npf, &c->d.pf_sb, &c->d.pf_eb, (long)g.ntchopdim * g.nrh, g.nrh, (long)g.npf * g.nrh, 1,
                       g.npf, g.nrh, 2 * SPLIT_KC)))
    return rc;
  // PF by degree: row (m, n) is contiguous over i; chunks of 16 nodes
  if ((rc = split_rows(c, c->d.pf, npf, &c->d.pf_sf, &c->d.pf_ef, (long)g.ntchopdim * g.npf, 1, g.nrh, 0, g.nrh, 1,
                       SPLIT_KC)))
    return rc;
  // V, D, T: row (t, m, n) contiguous over i
  return split_rows(c, c->d.vdt, nvdt, &c->d.vdt_s, &c->d.vdt_e, 3L * g.ntchopdim * g.nrchopdim, 1, g.nrh, 0, g.nrh, 1,
                    SPLIT_KC);
}

// -------------------------------------------------------------------------
// Backward: FFF (n,m,k) -> PFF (r,m,k).
// BE(i,c) = sum_j PF(i,2j) a(2j,c), BO(i,c) = sum_j PF(i,2j+1) a(2j+1,c);
// f(x_i) = BE+BO, f(x_{NR-1-i}) = BE-BO; LN*(-log(1-x)) added at m=0,k=0.
// Block tile: TI radial nodes x TC real columns (TC/2 complex k); the 4
// waves as (TI/32) x (128/TI), each wave 32 nodes (2 MFMA tiles) by
// TC TI/128 columns per parity.  K chunk: 16 pairs.  Every tile shape sums
// each output over the same sequence (chunks of 16 pairs, 4-deep MFMA steps
// in order), so all shapes give bitwise equal results.
// -------------------------------------------------------------------------
#define BW_JC 16

// LDS: one K-chunk stage of the table and data tiles (the next chunk waits
// in registers); after the K loop the same memory stages the output, one
// half (S = BE+BO, then D = BE-BO) at a time.  (64, 64): 33 KB per block,
// four blocks per CU.  Split mode: (hi, lo) pairs in place of the doubles,
// and the chunk's grid exponents per complex column and per table row.
template <bool SPLIT, int TI, int TC>
union RtranBwdLds {
  struct {
    typename LegElem<SPLIT>::T A[2][BW_JC][TI];  // [parity][j][i]
    typename LegElem<SPLIT>::T B[2][BW_JC][TC];  // [parity][j][col]
    int ce[SPLIT ? TC / 2 : 1];
    __align__(16) int re[SPLIT ? TI : 1];
  } k;
  double E[TC][TI + 1];  // [col][i]
};

// logical tiles of one backward product over the modes g.mlo .. g.mlo+g.mcnt-1
template <int TI, int TC>
__host__ __device__ inline int rtran_bwd_tiles(const KGrid &g, int nj) {
  return ((2 * g.nxchopdim + TC - 1) / TC) * ((g.nrh + TI - 1) / TI) * g.mcnt * nj;
}

// One block tile (logical id lid) of a backward product.  Logical order:
// radial tile fastest, then field, column tile, m: the radial tiles that
// read one data tile run one after another on one XCD (the data come from
// L2 after the first), and the mode's table tiles (all radial tiles of m)
// stay in that L2 for every field and column tile.
template <bool SPLIT, int TI, int TC>
__device__ __forceinline__ void rtran_bwd_tile(const KGrid &g, const typename LegElem<SPLIT>::T *__restrict__ pf,
                                               const TransformJobs &jobs, int nj, const double *__restrict__ lnx,
                                               const int *__restrict__ rowe, int lid, RtranBwdLds<SPLIT, TI, TC> &L) {
  typedef typename LegElem<SPLIT>::T TE;
  constexpr int WI = TI / 32, WC = 4 / WI, CW = TC / WC, SC = CW / 16;
  constexpr int QA = TI / 8, SA = 256 / TI, QB = TC / 16, QE = TI * TC / 512;
  static_assert((TI == 32 || TI == 64) && SC >= 1 && SC <= 2 && CW % 16 == 0, "tile shape");
  static_assert(QA * SA == 2 * BW_JC && QB * 8 == TC / 2 && QE * 512 == TI * TC, "staging covers the tile");
  const int nk = g.nxchopdim;
  const int ntiles_c = (2 * nk + TC - 1) / TC, ntiles_i = (g.nrh + TI - 1) / TI;
  const int ti = lid % ntiles_i, r1 = lid / ntiles_i;
  const int fj = r1 % nj, r2 = r1 / nj;
  const int tc = r2 % ntiles_c, mm = g.mlo + r2 / ntiles_c;
  const double2 *__restrict__ in = jobs.in[fj];
  const double *__restrict__ ln_in = jobs.ln[fj];
  double2 *__restrict__ out = jobs.out[fj];
  const int mv = mm * g.minc;
  const int nn = d_nrchops(g.nrchop, mv);
  const int i0 = ti * TI, c0 = tc * TC;
  // the thread id through an empty asm: in the grouped kernel's tile loop
  // the compiler then keeps no thread-dependent offsets of every tile shape
  // live across tiles (they would spill)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const int wi = wave % WI, wc = wave / WI;

  v4d acc[2][2][SC];  // [parity][sub i][sub c]
  v4f hh[2][2][SC], xx[2][2][SC];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < SC; b++) {
        acc[p][a][b] = (v4d){0, 0, 0, 0};
        hh[p][a][b] = xx[p][a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
      }

  const TE *pfm = pf + (size_t)mm * g.npf * g.nrh;
  const int njp = (nn + 1) / 2;  // (even, odd) pairs
  const int nch = (njp + BW_JC - 1) / BW_JC;
  const int nchr = (g.npf + 2 * BW_JC - 1) / (2 * BW_JC);  // chunks of a split table row
  // per-thread prefetch registers: QA table values, QB complex data values
  // (split: and one table row's grid exponent)
  TE ra[QA];
  double2 rb[QB];
  int rre = -1022;
  // 32-bit element offsets: lane part fixed, chunk and q parts uniform
  const int ta_i = i0 + (tid % TI), ta_n = tid / TI;          // table: i, n - 2 j0 - SA q
  const int tb_n = tid & 31, tb_k = (c0 >> 1) + (tid >> 5);   // data: n - 2 j0, k - 8 q
  const int kstride = g.nrchopdim * g.ntchopdim;
  const double2 *inm = in + (size_t)g.nrchopdim * mm;
  auto fetch = [&](int ch) {
    const int j0 = ch * BW_JC;
#pragma unroll
    for (int q = 0; q < QA; q++) {
      const int n = 2 * j0 + ta_n + SA * q;
      ra[q] = (n < nn && ta_i < g.nrh) ? pfm[n * g.nrh + ta_i] : TE{};
    }
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int n = 2 * j0 + tb_n, k = tb_k + 8 * q;
      rb[q] = (n < nn && k < nk) ? inm[n + kstride * k] : make_double2(0.0, 0.0);
    }
    if constexpr (SPLIT)
      if (tid < TI) rre = ta_i < g.nrh ? rowe[((size_t)mm * g.nrh + ta_i) * nchr + ch] : -1022;
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < QA; q++) {
      int idx = tid + 256 * q;
      int i = idx % TI, nl = idx / TI;
      L.k.A[nl & 1][nl >> 1][i ^ lsw(nl >> 1, 0)] = ra[q];
    }
    if constexpr (SPLIT)
      if (tid < TI) L.k.re[tid] = rre;
#pragma unroll
    for (int q = 0; q < QB; q++) {
      int idx = tid + 256 * q;
      int nl = idx & 31, kl = idx >> 5;
      const int sw = lsw(nl >> 1, nl & 1);
      if constexpr (SPLIT) {
        // the 32 rows of column kl in this chunk lie on 32 adjacent lanes
        const int e = grid_e(group_max<32>(max(bexp(rb[q].x), bexp(rb[q].y))));
        const double inv = inv_grid(e);
        if (nl == 0) L.k.ce[kl] = e;
        L.k.B[nl & 1][nl >> 1][(2 * kl) ^ sw] = split_hl(rb[q].x, inv);
        L.k.B[nl & 1][nl >> 1][(2 * kl + 1) ^ sw] = split_hl(rb[q].y, inv);
      } else {
        L.k.B[nl & 1][nl >> 1][(2 * kl) ^ sw] = rb[q].x;
        L.k.B[nl & 1][nl >> 1][(2 * kl + 1) ^ sw] = rb[q].y;
      }
    }
  };
  // the wave's column subtiles that hold data (the last tile of a mode is
  // partly padding); the others skip their products (wave-uniform)
  bool live[SC];
#pragma unroll
  for (int t = 0; t < SC; t++) live[t] = c0 + wc * CW + 16 * t < 2 * nk;
  if (nch > 0) fetch(0);
  for (int ch = 0; ch < nch; ch++) {
    __syncthreads();
    stash();
    __syncthreads();
    if (ch + 1 < nch) fetch(ch + 1);
    const int jleft = njp - ch * BW_JC;  // pairs of this chunk that hold data
#pragma unroll
    for (int ks = 0; ks < BW_JC; ks += 4) {
      if (ks >= jleft || !live[0]) break;  // zero padding (chunk tail, or all the wave's columns)
      const int kr = ks + (lane >> 4);
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const int sa = lsw(kr, 0), sb = lsw(kr, p);
        const TE a0 = L.k.A[p][kr][(wi * 32 + (lane & 15)) ^ sa];
        const TE a1 = L.k.A[p][kr][(wi * 32 + 16 + (lane & 15)) ^ sa];
#pragma unroll
        for (int t = 0; t < SC; t++) {
          if (t > 0 && !live[t]) break;
          const TE bt = L.k.B[p][kr][(wc * CW + 16 * t + (lane & 15)) ^ sb];
          if constexpr (SPLIT) {
            mfma_split(a0, bt, hh[p][0][t], xx[p][0][t]);
            mfma_split(a1, bt, hh[p][1][t], xx[p][1][t]);
          } else {
            acc[p][0][t] = mfma16x16x4(a0, bt, acc[p][0][t]);
            acc[p][1][t] = mfma16x16x4(a1, bt, acc[p][1][t]);
          }
        }
      }
    }
    if constexpr (SPLIT) {
#pragma unroll
      for (int t = 0; t < SC; t++) {
        const int ec = L.k.ce[(wc * CW + 16 * t + (lane & 15)) >> 1];
#pragma unroll
        for (int a = 0; a < 2; a++) {
          const int4 er = *(const int4 *)&L.k.re[wi * 32 + a * 16 + 4 * (lane >> 4)];
#pragma unroll
          for (int p = 0; p < 2; p++) fold_split(acc[p][a][t], hh[p][a][t], xx[p][a][t], er, ec);
        }
      }
    }
  }
  // epilogue through LDS as [col][i]: S = BE+BO to f(x_i), then D = BE-BO
  // to f(x_{NR-1-i})
  const double lnv = (mm == 0) ? ln_in[0] : 0.0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < SC; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          int il = wi * 32 + a * 16 + acc_row<SPLIT>(lane, v);
          int cl = wc * CW + b * 16 + (lane & 15);
          double e = acc[0][a][b][v], o = acc[1][a][b][v];
          L.E[cl][il] = h == 0 ? e + o : e - o;
        }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QE; q++) {
      int idx = tid + 256 * q;  // i fastest
      int il = idx % TI, kl = idx / TI;
      int ig = i0 + il, k = (c0 >> 1) + kl;
      if (ig < g.nrh && k < nk) {
        size_t col = (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * k);
        double2 v = make_double2(L.E[2 * kl][il], L.E[2 * kl + 1][il]);
        const int ir = h == 0 ? ig : g.nr - 1 - ig;
        if (mm == 0 && k == 0 && lnv != 0.0) v.x = v.x + lnv * lnx[ir];
        out[col + ir] = v;
      }
    }
  }
  // the padding rows r = NR..NDIMR-1 of the block's columns are zero (CHOPDO
  // of the PFF result)
  if (ti == 0) {
    const int pad = g.ndimr - g.nr;
    for (int idx = tid; idx < pad * (TC / 2); idx += 256) {
      const int kl = idx / pad, k = (c0 >> 1) + kl;
      if (k < nk) out[(size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * k) + g.nr + (idx - kl * pad)] = make_double2(0.0, 0.0);
    }
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_rtran_bwd(KGrid g, const double *__restrict__ pf, TransformJobs jobs, int nj, const double *__restrict__ lnx) {
  __shared__ RtranBwdLds<false, 64, 64> L;
  const int lid = xcd_logical(rtran_bwd_tiles<64, 64>(g, nj));
  if (lid < 0) return;
  rtran_bwd_tile<false, 64, 64>(g, pf, jobs, nj, lnx, nullptr, lid, L);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_rtran_bwd_split(KGrid g, const float2 *__restrict__ pf, TransformJobs jobs, int nj, const double *__restrict__ lnx,
                  const int *__restrict__ rowe) {
  __shared__ RtranBwdLds<true, 64, 64> L;
  const int lid = xcd_logical(rtran_bwd_tiles<64, 64>(g, nj));
  if (lid < 0) return;
  rtran_bwd_tile<true, 64, 64>(g, pf, jobs, nj, lnx, rowe, lid, L);
}

// ---- grouped backward products: any number of (run, mode range) jobs in
// one persistent launch, each job with its own tile shape.
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t BWD_LDS_BYTES =
    cmax(sizeof(RtranBwdLds<false, 64, 64>), cmax(sizeof(RtranBwdLds<false, 32, 64>), sizeof(RtranBwdLds<false, 64, 32>)));
template <int TI, int TC>
__device__ __forceinline__ void bwd_job_tile(const BwdJob &j, int lid, unsigned char *lds) {
  rtran_bwd_tile<false, TI, TC>(j.g, j.pf, j.jobs, j.nj, j.lnx, nullptr, lid,
                                *reinterpret_cast<RtranBwdLds<false, TI, TC> *>(lds));
}

// Logical tiles 0 .. total-1, job d holding tiles b0[d] .. b0[d+1]-1
// (b0[njobs] = total).  The tiles split into 8 contiguous parts, one per
// XCD; the blocks of one XCD (linear ids = x mod 8; gridDim.x a multiple of
// 8) walk its part with stride gridDim.x / 8, a barrier between tiles.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_rtran_bwd_grouped(const BwdJob *__restrict__ jb, const int *__restrict__ b0, int njobs, int total) {
  __shared__ __align__(16) unsigned char lds[BWD_LDS_BYTES];
  const int per = (total + 7) >> 3;
  const int x = blockIdx.x & 7, stride = gridDim.x >> 3;
  const int lo = x * per, hi = min(total, lo + per);
  const int first = lo + (int)(blockIdx.x >> 3);
  // the job of the block's first tile: the last d with b0[d] <= first
  int d = 0;
  for (int span = njobs; span > 1;) {
    const int h = span >> 1;
    if (b0[d + h] <= first) d += h;
    span -= h;
  }
  for (int lid = first; lid < hi; lid += stride) {
    while (d + 1 < njobs && b0[d + 1] <= lid) d++;  // tiles of a block only increase
    const BwdJob &j = jb[d];
    const int t = lid - b0[d];
    switch (j.shape) {
      case BWD_TILE_32x64: bwd_job_tile<32, 64>(j, t, lds); break;
      case BWD_TILE_64x32: bwd_job_tile<64, 32>(j, t, lds); break;
      default: bwd_job_tile<64, 64>(j, t, lds); break;
    }
    __syncthreads();
  }
}

int launch_rtran_bwd_n(mlg_ctx *c, int n, const double2 *const *in, const double *const *ln_in, double2 *const *out) {
  if (n < 1 || n > MAX_TRANSFORM_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  TransformJobs jobs;
  for (int i = 0; i < n; i++) {
    jobs.in[i] = in[i];
    jobs.ln[i] = ln_in[i];
    jobs.out[i] = out[i];
  }
  // 2 NR sum_m NRCHOPS(m) NXCHOPDIM flops per field (real table x complex data, folded)
  const double per_k = n * 2.0 * g.nr * g.nxchopdim;
  if (!c->leg_split) {
    if (c->rec) return sweep_defer_bwd(c, g, jobs, n);
    KScope ks(c, K_RTRAN_BWD, per_k * (double)sum_nrchops(c));
    hipLaunchKernelGGL(k_rtran_bwd, dim3(xcd_grid(rtran_bwd_tiles<64, 64>(g, n))), dim3(256), 0, c->stream, g, c->d.pf,
                       jobs, n, c->d.lnx);
    MLG_CHECK_HIP(hipGetLastError());
    return 0;
  }
  KGrid g0, g1;
  split_modes(c, g, g0, g1);
  {
    KScope ks(c, K_RTRAN_BWD, per_k * split_fp64_sum(g0, 1));
    hipLaunchKernelGGL(k_rtran_bwd, dim3(xcd_grid(rtran_bwd_tiles<64, 64>(g0, n))), dim3(256), 0, c->stream, g0, c->d.pf,
                       jobs, n, c->d.lnx);
    MLG_CHECK_HIP(hipGetLastError());
  }
  if (g1.mcnt > 0) {
    KScope ks(c, K_RTRAN_BWD, per_k * ((double)sum_nrchops(c) - split_fp64_sum(g0, 1)));
    hipLaunchKernelGGL(k_rtran_bwd_split, dim3(xcd_grid(rtran_bwd_tiles<64, 64>(g1, n))), dim3(256), 0, c->stream, g1,
                       c->d.pf_sb, jobs, n, c->d.lnx, c->d.pf_eb);
    MLG_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

int launch_rtran_bwd(mlg_ctx *c, const double2 *in, const double *ln_in, double2 *out) {
  return launch_rtran_bwd_n(c, 1, &in, &ln_in, &out);
}

int bwd_job_tiles(const BwdJob &j) {
  switch (j.shape) {
    case BWD_TILE_32x64: return rtran_bwd_tiles<32, 64>(j.g, j.nj);
    case BWD_TILE_64x32: return rtran_bwd_tiles<64, 32>(j.g, j.nj);
    default: return rtran_bwd_tiles<64, 64>(j.g, j.nj);
  }
}

// three resident blocks per CU (3 waves per SIMD at 256 threads), rounded
// to the 8 XCDs; no more blocks than tiles
int bwd_grouped_grid(int total, int cus) {
  const int cap = (3 * cus + 7) & ~7;
  return std::max(8, std::min(xcd_grid(total), cap));
}

void *bwd_grouped_kernel() { return (void *)k_rtran_bwd_grouped; }

int launch_rtran_bwd_grouped(hipStream_t st, const BwdJob *jobs, const int *b0, int njobs, int total, int grid) {
  if (njobs < 1 || grid < 8 || (grid & 7)) return MLG_E_ARG;
  hipLaunchKernelGGL(k_rtran_bwd_grouped, dim3(grid), dim3(256), 0, st, jobs, b0, njobs, total);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// -------------------------------------------------------------------------
// Forward: PFF (r,m,k) -> FFF (n,m,k).
// a_n = sum_i PF(i,n) F_p(i,c), F_e = (f(x_i)+f(x_{NR-1-i})) w_i,
// F_o = (f(x_i)-f(x_{NR-1-i})) w_i, p = parity of n.
// Block tile: 64 rows n (32 even + 32 odd) x 64 real columns; wave w:
// parity w&1, column half w>>1 (2 x 2 MFMA tiles).  K chunk: 16 nodes,
// staged in LDS while the next chunk is fetched into registers.  Rows n in
// [NRCHOPS(m), NRCHOPDIM) and modes m >= NTCHOP of the output are written as
// zeros (the CHOPDO of the result).
// -------------------------------------------------------------------------
#define FW_TN 64
#define FW_TC 64
#define FW_IC 16

template <bool SPLIT>
union RtranFwdLds {
  struct {
    typename LegElem<SPLIT>::T A[2][FW_IC][FW_TN / 2];  // [parity][i][n-local within parity]
    typename LegElem<SPLIT>::T B[2][FW_IC][FW_TC];      // [fold parity][i][col]
    int ce[2][SPLIT ? FW_TC / 2 : 1];                   // split: [fold parity][complex col]
    __align__(16) int re[2][SPLIT ? FW_TN / 2 : 1];     // split: [parity][row within parity]
  } k;
  double C[FW_TC][FW_TN + 1];  // [col][n-local]
};

template <bool SPLIT>
__device__ __forceinline__ void rtran_fwd_body(const KGrid &g, const typename LegElem<SPLIT>::T *__restrict__ pf,
                                               const double *__restrict__ w, const double2 *__restrict__ in,
                                               double2 *__restrict__ out, double *zl, const int *__restrict__ rowe,
                                               RtranFwdLds<SPLIT> &L) {
  typedef typename LegElem<SPLIT>::T TE;
  if (blockIdx.x == 0 && threadIdx.x == 0 && zl) zl[0] = 0.0;  // the result's LN
  const int nk = g.nxchopdim;
  const int ntc = (2 * nk + FW_TC - 1) / FW_TC, ntn = (g.nrchopdim + FW_TN - 1) / FW_TN;
  const int lid = xcd_logical(ntc * ntn * g.mcnt);
  if (lid < 0) return;
  const int ml = lid / (ntc * ntn), rem = lid - ml * (ntc * ntn), mm = g.mlo + ml;
  const int tn = rem / ntc, tc = rem - tn * ntc;
  const int mv = mm * g.minc;
  const int nn = mm < g.ntchop ? d_nrchops(g.nrchop, mv) : 0;
  const int n0 = tn * FW_TN, c0 = tc * FW_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n0 >= nn) {  // a tile of the truncated rows: zero (CHOPDO)
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q, n = n0 + (idx & 63), k = (c0 >> 1) + (idx >> 6);
      if (n < g.nrchopdim && k < nk)
        out[(size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k)] = make_double2(0.0, 0.0);
    }
    return;
  }
  const int par = wave & 1, ch = wave >> 1;

  v4d acc[2][2] = {{(v4d){0, 0, 0, 0}, (v4d){0, 0, 0, 0}}, {(v4d){0, 0, 0, 0}, (v4d){0, 0, 0, 0}}};
  v4f hh[2][2], xx[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) hh[a][b] = xx[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
  const TE *pfm = pf + (size_t)mm * g.npf * g.nrh;
  const double2 *inm = in + (size_t)g.ndimr * mm;
  const int kstride = g.ndimr * g.ntchopdim;
  // lane parts of the fetch indices (table: i fastest; data: i fastest)
  const int il = tid & 15, nl0 = tid >> 4, kl0 = tid >> 4;
  TE ra[4];
  double2 rp[2], rq[2];
  double rw = 0.0;
  const int nchr = (g.nrh + FW_IC - 1) / FW_IC;  // chunks of a split table row
  int rre = -1022;
  auto fetch = [&](int ib) {
    const int i = ib + il;
    if constexpr (SPLIT)
      if (tid < FW_TN) rre = n0 + tid < nn ? rowe[((size_t)mm * g.npf + n0 + tid) * nchr + ib / FW_IC] : -1022;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = n0 + nl0 + 16 * q;
      ra[q] = (n < nn && i < g.nrh) ? pfm[n * g.nrh + i] : TE{};
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int k = (c0 >> 1) + kl0 + 16 * q;
      rp[q] = rq[q] = make_double2(0.0, 0.0);
      if (i < g.nrh && k < nk) {
        rp[q] = inm[kstride * k + i];
        rq[q] = inm[kstride * k + (g.nr - 1 - i)];
      }
    }
    rw = (i < g.nrh) ? w[i] : 0.0;
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int nl = nl0 + 16 * q;
      L.k.A[nl & 1][il][(nl >> 1) ^ lsw(il, nl & 1)] = ra[q];
    }
    if constexpr (SPLIT)
      if (tid < FW_TN) L.k.re[tid & 1][tid >> 1] = rre;
    const int sw = lsw(il, 0);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int kl = kl0 + 16 * q;
      const double2 p = rp[q], r = rq[q];
      const double ex = (p.x + r.x) * rw, ey = (p.y + r.y) * rw;
      const double ox = (p.x - r.x) * rw, oy = (p.y - r.y) * rw;
      if constexpr (SPLIT) {
        // the chunk's 16 nodes of column kl lie on 16 adjacent lanes
        const int ee = grid_e(group_max<16>(max(bexp(ex), bexp(ey))));
        const int eo = grid_e(group_max<16>(max(bexp(ox), bexp(oy))));
        if (il == 0) {
          L.k.ce[0][kl] = ee;
          L.k.ce[1][kl] = eo;
        }
        const double ie = inv_grid(ee), io = inv_grid(eo);
        L.k.B[0][il][(2 * kl) ^ sw] = split_hl(ex, ie);
        L.k.B[0][il][(2 * kl + 1) ^ sw] = split_hl(ey, ie);
        L.k.B[1][il][(2 * kl) ^ sw] = split_hl(ox, io);
        L.k.B[1][il][(2 * kl + 1) ^ sw] = split_hl(oy, io);
      } else {
        L.k.B[0][il][(2 * kl) ^ sw] = ex;
        L.k.B[0][il][(2 * kl + 1) ^ sw] = ey;
        L.k.B[1][il][(2 * kl) ^ sw] = ox;
        L.k.B[1][il][(2 * kl + 1) ^ sw] = oy;
      }
    }
  };
  // subtiles that hold data: rows n0 + 32 a + par + 2 l, columns c0 + 32 ch + 16 b + l
  const bool ra0 = n0 + par < nn, ra1 = n0 + 32 + par < nn;
  const bool rc0 = c0 + ch * 32 < 2 * nk, rc1 = c0 + ch * 32 + 16 < 2 * nk;
  fetch(0);
  for (int ib = 0; ib < g.nrh; ib += FW_IC) {
    __syncthreads();
    stash();
    __syncthreads();
    if (ib + FW_IC < g.nrh) fetch(ib + FW_IC);
    if (!(ra0 && rc0)) continue;  // all the wave's rows or columns are padding
#pragma unroll
    for (int ks = 0; ks < FW_IC; ks += 4) {
      const int kr = ks + (lane >> 4), sw = lsw(kr, 0), swa = lsw(kr, par);
      const TE a0 = L.k.A[par][kr][(lane & 15) ^ swa];
      const TE a1 = L.k.A[par][kr][(16 + (lane & 15)) ^ swa];
      const TE b0 = L.k.B[par][kr][(ch * 32 + (lane & 15)) ^ sw];
      const TE b1 = L.k.B[par][kr][(ch * 32 + 16 + (lane & 15)) ^ sw];
      if constexpr (SPLIT) {
        mfma_split(a0, b0, hh[0][0], xx[0][0]);
        if (rc1) mfma_split(a0, b1, hh[0][1], xx[0][1]);
        if (ra1) {
          mfma_split(a1, b0, hh[1][0], xx[1][0]);
          if (rc1) mfma_split(a1, b1, hh[1][1], xx[1][1]);
        }
      } else {
        acc[0][0] = mfma16x16x4(a0, b0, acc[0][0]);
        if (rc1) acc[0][1] = mfma16x16x4(a0, b1, acc[0][1]);
        if (ra1) {
          acc[1][0] = mfma16x16x4(a1, b0, acc[1][0]);
          if (rc1) acc[1][1] = mfma16x16x4(a1, b1, acc[1][1]);
        }
      }
    }
    if constexpr (SPLIT) {
      const int ec0 = L.k.ce[par][(ch * 32 + (lane & 15)) >> 1];
      const int ec1 = L.k.ce[par][(ch * 32 + 16 + (lane & 15)) >> 1];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int4 er = *(const int4 *)&L.k.re[par][16 * a + 4 * (lane >> 4)];
        fold_split(acc[a][0], hh[a][0], xx[a][0], er, ec0);
        fold_split(acc[a][1], hh[a][1], xx[a][1], er, ec1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rl = 16 * a + acc_row<SPLIT>(lane, v);  // row within the parity tile
      const int nl = 2 * rl + par;                        // n-local
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int cl = ch * 32 + b * 16 + (lane & 15);
        L.C[cl][nl] = acc[a][b][v];
      }
    }
  __syncthreads();
  for (int q = 0; q < 8; q++) {
    const int idx = tid + 256 * q;  // 0..2047, n fastest
    const int nl = idx & 63, kl = idx >> 6;
    const int n = n0 + nl, k = (c0 >> 1) + kl;
    if (n < g.nrchopdim && k < nk) {
      double2 v = (n < nn) ? make_double2(L.C[2 * kl][nl], L.C[2 * kl + 1][nl]) : make_double2(0.0, 0.0);
      out[(size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k)] = v;
    }
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_rtran_fwd(KGrid g, const double *__restrict__ pf, const double *__restrict__ w, const double2 *__restrict__ in,
            double2 *__restrict__ out, double *zl) {
  __shared__ RtranFwdLds<false> L;
  rtran_fwd_body<false>(g, pf, w, in, out, zl, nullptr, L);
}

// n fields in one launch (grid row = field), FP64 products
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_rtran_fwd_n(KGrid g, const double *__restrict__ pf, const double *__restrict__ w, TransformJobs jobs) {
  __shared__ RtranFwdLds<false> L;
  rtran_fwd_body<false>(g, pf, w, jobs.in[blockIdx.y], jobs.out[blockIdx.y], nullptr, nullptr, L);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_rtran_fwd_split(KGrid g, const float2 *__restrict__ pf, const double *__restrict__ w, const double2 *__restrict__ in,
                  double2 *__restrict__ out, double *zl, const int *__restrict__ rowe) {
  __shared__ RtranFwdLds<true> L;
  rtran_fwd_body<true>(g, pf, w, in, out, zl, rowe, L);
}

int launch_rtran_fwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out, const double *w) {
  if (n < 1 || n > MAX_TRANSFORM_JOBS || c->leg_split) return MLG_E_ARG;
  KGrid g = kgrid(c);
  int ntn = (g.nrchopdim + FW_TN - 1) / FW_TN, ntc = (2 * g.nxchopdim + FW_TC - 1) / FW_TC;
  TransformJobs jobs;
  for (int i = 0; i < n; i++) {
    jobs.in[i] = in[i];
    jobs.ln[i] = nullptr;
    jobs.out[i] = out[i];
  }
  KScope ks(c, K_RTRAN_FWD, n * 2.0 * g.nr * g.nxchopdim * (double)sum_nrchops(c));
  hipLaunchKernelGGL(k_rtran_fwd_n, dim3(xcd_grid(ntn * ntc * g.mcnt), n), dim3(256), 0, c->stream, g, c->d.pf,
                     w ? w : c->d.w, jobs);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_rtran_fwd(mlg_ctx *c, const double2 *in, double2 *out, double *zl, const double *w) {
  KGrid g = kgrid(c);
  int ntn = (g.nrchopdim + FW_TN - 1) / FW_TN, ntc = (2 * g.nxchopdim + FW_TC - 1) / FW_TC;
  const double per_k = 2.0 * g.nr * g.nxchopdim;
  if (w && c->leg_split) return MLG_E_ARG;
  if (!c->leg_split) {
    KScope ks(c, K_RTRAN_FWD, per_k * (double)sum_nrchops(c));
    hipLaunchKernelGGL(k_rtran_fwd, dim3(xcd_grid(ntn * ntc * g.mcnt)), dim3(256), 0, c->stream, g, c->d.pf, w ? w : c->d.w,
                       in, out, zl);
    MLG_CHECK_HIP(hipGetLastError());
    return 0;
  }
  KGrid g0, g1;
  split_modes(c, g, g0, g1);
  {
    KScope ks(c, K_RTRAN_FWD, per_k * split_fp64_sum(g0, 1));
    hipLaunchKernelGGL(k_rtran_fwd, dim3(xcd_grid(ntn * ntc * g0.mcnt)), dim3(256), 0, c->stream, g0, c->d.pf, c->d.w, in,
                       out, zl);
    MLG_CHECK_HIP(hipGetLastError());
  }
  if (g1.mcnt > 0) {
    KScope ks(c, K_RTRAN_FWD, per_k * ((double)sum_nrchops(c) - split_fp64_sum(g0, 1)));
    hipLaunchKernelGGL(k_rtran_fwd_split, dim3(xcd_grid(ntn * ntc * g1.mcnt)), dim3(256), 0, c->stream, g1, c->d.pf_sf,
                       c->d.w, in, out, nullptr, c->d.pf_ef);
    MLG_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

// -------------------------------------------------------------------------
// PROJECT products: five parity-folded products per m,
//   W1 = EO(V, RUR), W2 = OE(D, RUP), W3 = OE(D, RUR), W4 = EO(V, RUP),
//   W5 = EO(T, UZ),
// psi = -i m W1 - W2 ; del^2 chi = i k W3 + m k W4 - W5.
// EO: even n uses the even fold (u(x_i)+u(x_{NR-1-i})), odd n the odd fold;
// OE swaps them.  Folds here carry no weight (it is inside V, D, T).
// -------------------------------------------------------------------------
#define PJ_TN 64
#define PJ_TC 32
#define PJ_IC 16

// Split mode: 32 rows per block (one 16-row subtile per wave and parity),
// so the FP32 accumulators fit beside the FP64 ones.
template <bool SPLIT>
union ProjectLds {
  struct {
    typename LegElem<SPLIT>::T T[3][2][PJ_IC][PJ_TN / 2];  // [V,D,T][parity][i][n-local within parity]
    typename LegElem<SPLIT>::T F[3][2][PJ_IC][PJ_TC];      // [field][fold parity][i][col]
    int ce[3][2][SPLIT ? PJ_TC / 2 : 1];                   // split: [field][fold parity][complex col]
    __align__(16) int re[3][2][SPLIT ? 16 : 1];            // split: [table][parity][row within parity]
  } k;
  struct {
    double P[PJ_TC / 2][PJ_TN + 1][2];  // psi   [complex k][n][re,im]
    double C[PJ_TC / 2][PJ_TN + 1][2];  // del^2 chi
  } e;
};

// psi = -i m W1 - W2, del^2 chi = i k W3 + m k W4 - W5 from the lane's five
// accumulators, staged in LDS, then stored with rows
// n >= NRCHOPS(m) zero; rows per block TN.
template <bool SPLIT, int TN, int NA>
__device__ __forceinline__ void project_epilogue(const KGrid &g, ProjectLds<SPLIT> &L, v4d (&acc)[NA][5],
                                                 const double *__restrict__ ak, int mm,
                                                 int mv, int nn, int n0, int c0, int par, int ch, double2 *__restrict__ psi,
                                                 double2 *__restrict__ chi) {
  const int nk = g.nxchopdim, tid = threadIdx.x, lane = tid & 63;
  const int cl = ch * 16 + (lane & 15), kl = cl >> 1, isim = cl & 1;
  const int kg = (c0 >> 1) + kl;
  const double m = (double)mv;
  const double kv = (kg < nk) ? ak[mm + (size_t)g.ntchopdim * kg] : 0.0, mk = m * kv;
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rl = acc_row<SPLIT>(lane, v) + 16 * a;
      const int nl = 2 * rl + par;
      const double w1 = acc[a][0][v], w2 = acc[a][1][v], w3 = acc[a][2][v], w4 = acc[a][3][v], w5 = acc[a][4][v];
      const double w1p = __shfl_xor(w1, 1, 64), w3p = __shfl_xor(w3, 1, 64);
      double ps, cs;
      if (!isim) {  // real parts
        ps = m * w1p - w2;
        cs = (-kv * w3p + mk * w4) - w5;
      } else {      // imaginary parts
        ps = -m * w1p - w2;
        cs = (kv * w3p + mk * w4) - w5;
      }
      L.e.P[kl][nl][isim] = ps;
      L.e.C[kl][nl][isim] = cs;
    }
  __syncthreads();
  for (int q = 0; q < TN / 16; q++) {
    int idx = tid + 256 * q;  // n fastest, 16 complex k
    int nl = idx % TN, kq = idx / TN;
    int n = n0 + nl, k = (c0 >> 1) + kq;
    if (n < g.nrchopdim && k < nk) {
      size_t o = (size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
      if (n < nn) {
        psi[o] = make_double2(L.e.P[kq][nl][0], L.e.P[kq][nl][1]);
        chi[o] = make_double2(L.e.C[kq][nl][0], L.e.C[kq][nl][1]);
      } else {
        psi[o] = make_double2(0.0, 0.0);
        chi[o] = make_double2(0.0, 0.0);
      }
    }
  }
}

// a tile of truncated rows (n0 >= NRCHOPS(m)): zeros (CHOPDO)
template <int TN>
__device__ __forceinline__ void project_zero_tile(const KGrid &g, int mm, int n0, int c0, double2 *__restrict__ psi,
                                                  double2 *__restrict__ chi) {
  const int tid = threadIdx.x;
  for (int q = 0; q < TN / 16; q++) {
    const int idx = tid + 256 * q, n = n0 + idx % TN, k = (c0 >> 1) + idx / TN;
    if (n < g.nrchopdim && k < g.nxchopdim) {
      const size_t o = (size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
      psi[o] = make_double2(0.0, 0.0);
      chi[o] = make_double2(0.0, 0.0);
    }
  }
}

// Block tile: 64 rows n (32 even + 32 odd; split: 32) x 32 real columns (16
// complex k); wave w: parity w&1, column half w>>1, two (split: one) 16-row
// subtiles.  The K loop (Gauss nodes, chunks of 16) stages one chunk in LDS
// while the next one is being fetched into registers.  The epilogue forms
// psi and del^2 chi in registers, exchanging re/im partners with a lane
// shuffle.
template <bool SPLIT>
__device__ __forceinline__ void project_body(const KGrid &g, const typename LegElem<SPLIT>::T *__restrict__ vdt,
                                             const double *__restrict__ ak, const double2 *__restrict__ rur,
                                             const double2 *__restrict__ rup, const double2 *__restrict__ uz,
                                             double2 *__restrict__ psi, double2 *__restrict__ chi,
                                             const int *__restrict__ rowe, ProjectLds<SPLIT> &L) {
  typedef typename LegElem<SPLIT>::T TE;
  constexpr int TN = SPLIT ? 32 : PJ_TN, NA = SPLIT ? 1 : 2, NQT = 3 * TN / 16;
  const int nk = g.nxchopdim;
  const int ntc = (2 * nk + PJ_TC - 1) / PJ_TC, ntn = (g.nrchopdim + TN - 1) / TN;
  const int lid = xcd_logical(ntc * ntn * g.mcnt);
  if (lid < 0) return;
  const int ml = lid / (ntc * ntn), rem = lid - ml * (ntc * ntn), mm = g.mlo + ml;
  const int tn = rem / ntc, tc = rem - tn * ntc;
  const int mv = mm * g.minc;
  const int nn = mm < g.ntchop ? d_nrchops(g.nrchop, mv) : 0;
  const int n0 = tn * TN, c0 = tc * PJ_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n0 >= nn) {
    project_zero_tile<TN>(g, mm, n0, c0, psi, chi);
    return;
  }
  const int par = wave & 1, ch = wave >> 1;
  const size_t tstride = (size_t)g.ntchopdim * g.nrchopdim * g.nrh;

  v4d acc[NA][5];
  v4f hh[NA][5], xx[NA][5];
  for (int a = 0; a < NA; a++)
    for (int q = 0; q < 5; q++) {
      acc[a][q] = (v4d){0, 0, 0, 0};
      hh[a][q] = xx[a][q] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
  const double2 *src[3] = {rur, rup, uz};
  TE rt[NQT];
  double2 rp[3], rq[3];
  const int nchr = (g.nrh + PJ_IC - 1) / PJ_IC;  // chunks of a split table row
  int rre = -1022;
  // 32-bit element offsets from per-m bases
  const TE *vdm = vdt + (size_t)mm * g.nrchopdim * g.nrh;
  const int its = (int)tstride, kstride = g.ndimr * g.ntchopdim;
  const size_t moff = (size_t)g.ndimr * mm;
  auto fetch = [&](int ib) {
    if constexpr (SPLIT)
      if (tid < 3 * TN) {  // [table 3][n TN]
        const int t = tid / TN, n = n0 + tid % TN;
        rre = n < nn ? rowe[(((size_t)t * g.ntchopdim + mm) * g.nrchopdim + n) * nchr + ib / PJ_IC] : -1022;
      }
#pragma unroll
    for (int q = 0; q < NQT; q++) {
      int idx = tid + 256 * q;  // [table 3][n TN][i 16]
      int il = idx & 15, nl = (idx >> 4) % TN, t = idx / (16 * TN);
      int n = n0 + nl, i = ib + il;
      rt[q] = (n < nn && i < g.nrh) ? vdm[t * its + n * g.nrh + i] : TE{};
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int idx = tid + 256 * q;  // [field 3][k 16][i 16]
      int il = idx & 15, kl = (idx >> 4) & 15, f = idx >> 8;
      int i = ib + il, k = (c0 >> 1) + kl;
      rp[q] = rq[q] = make_double2(0.0, 0.0);
      if (i < g.nrh && k < nk) {
        const double2 *col = src[f] + moff;
        rp[q] = col[kstride * k + i];
        rq[q] = col[kstride * k + (g.nr - 1 - i)];
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < NQT; q++) {
      int idx = tid + 256 * q;
      int il = idx & 15, nl = (idx >> 4) % TN, t = idx / (16 * TN);
      L.k.T[t][nl & 1][il][(nl >> 1) ^ lsw(il, nl & 1)] = rt[q];
    }
    if constexpr (SPLIT)
      if (tid < 3 * TN) L.k.re[tid / TN][(tid % TN) & 1][(tid % TN) >> 1] = rre;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int idx = tid + 256 * q;
      int il = idx & 15, kl = (idx >> 4) & 15, f = idx >> 8;
      double2 p = rp[q], r = rq[q];
      const int sw = lsw(il, 0);
      const double ex = p.x + r.x, ey = p.y + r.y, ox = p.x - r.x, oy = p.y - r.y;
      if constexpr (SPLIT) {
        // the chunk's 16 nodes of (field, column) lie on 16 adjacent lanes
        const int ee = grid_e(group_max<16>(max(bexp(ex), bexp(ey))));
        const int eo = grid_e(group_max<16>(max(bexp(ox), bexp(oy))));
        if (il == 0) {
          L.k.ce[f][0][kl] = ee;
          L.k.ce[f][1][kl] = eo;
        }
        const double ie = inv_grid(ee), io = inv_grid(eo);
        L.k.F[f][0][il][(2 * kl) ^ sw] = split_hl(ex, ie);
        L.k.F[f][0][il][(2 * kl + 1) ^ sw] = split_hl(ey, ie);
        L.k.F[f][1][il][(2 * kl) ^ sw] = split_hl(ox, io);
        L.k.F[f][1][il][(2 * kl + 1) ^ sw] = split_hl(oy, io);
      } else {
        L.k.F[f][0][il][(2 * kl) ^ sw] = ex;
        L.k.F[f][0][il][(2 * kl + 1) ^ sw] = ey;
        L.k.F[f][1][il][(2 * kl) ^ sw] = ox;
        L.k.F[f][1][il][(2 * kl + 1) ^ sw] = oy;
      }
    }
  };
  const int eo = par, oe = par ^ 1;
  // subtiles that hold data: rows n0 + 32 a + par + 2 l, columns c0 + 16 ch + l
  const bool pa0 = n0 + par < nn, pa1 = NA > 1 && n0 + 32 + par < nn, pc0 = c0 + ch * 16 < 2 * nk;
  fetch(0);
  for (int ib = 0; ib < g.nrh; ib += PJ_IC) {
    __syncthreads();
    stash();
    __syncthreads();
    if (ib + PJ_IC < g.nrh) fetch(ib + PJ_IC);
    if (!(pa0 && pc0)) continue;  // all the wave's rows or columns are padding
#pragma unroll
    for (int ks = 0; ks < PJ_IC; ks += 4) {
      const int kr = ks + (lane >> 4), cl = (ch * 16 + (lane & 15)) ^ lsw(kr, 0), st = lsw(kr, par);
      const TE f0e = L.k.F[0][eo][kr][cl], f0o = L.k.F[0][oe][kr][cl];
      const TE f1e = L.k.F[1][eo][kr][cl], f1o = L.k.F[1][oe][kr][cl];
      const TE f2e = L.k.F[2][eo][kr][cl];
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (a == 1 && !pa1) break;
        const int rl = (16 * a + (lane & 15)) ^ st;
        const TE av = L.k.T[0][par][kr][rl], ad = L.k.T[1][par][kr][rl], at = L.k.T[2][par][kr][rl];
        if constexpr (SPLIT) {
          mfma_split(av, f0e, hh[a][0], xx[a][0]);
          mfma_split(ad, f1o, hh[a][1], xx[a][1]);
          mfma_split(ad, f0o, hh[a][2], xx[a][2]);
          mfma_split(av, f1e, hh[a][3], xx[a][3]);
          mfma_split(at, f2e, hh[a][4], xx[a][4]);
        } else {
          acc[a][0] = mfma16x16x4(av, f0e, acc[a][0]);
          acc[a][1] = mfma16x16x4(ad, f1o, acc[a][1]);
          acc[a][2] = mfma16x16x4(ad, f0o, acc[a][2]);
          acc[a][3] = mfma16x16x4(av, f1e, acc[a][3]);
          acc[a][4] = mfma16x16x4(at, f2e, acc[a][4]);
        }
      }
    }
    if constexpr (SPLIT) {
      // one 16-row subtile per wave: rows 4 (lane / 16) + v of its parity
      const int kl = (ch * 16 + (lane & 15)) >> 1, rb = 4 * (lane >> 4);
      const int4 ev = *(const int4 *)&L.k.re[0][par][rb], ed = *(const int4 *)&L.k.re[1][par][rb];
      const int4 et = *(const int4 *)&L.k.re[2][par][rb];
      fold_split(acc[0][0], hh[0][0], xx[0][0], ev, L.k.ce[0][eo][kl]);
      fold_split(acc[0][1], hh[0][1], xx[0][1], ed, L.k.ce[1][oe][kl]);
      fold_split(acc[0][2], hh[0][2], xx[0][2], ed, L.k.ce[0][oe][kl]);
      fold_split(acc[0][3], hh[0][3], xx[0][3], ev, L.k.ce[1][eo][kl]);
      fold_split(acc[0][4], hh[0][4], xx[0][4], et, L.k.ce[2][eo][kl]);
    }
  }
  __syncthreads();
  project_epilogue<SPLIT, TN, NA>(g, L, acc, ak, mm, mv, nn, n0, c0, par, ch, psi, chi);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_project(KGrid g, const double *__restrict__ vdt, const double *__restrict__ ak, const double2 *__restrict__ rur,
          const double2 *__restrict__ rup, const double2 *__restrict__ uz, double2 *__restrict__ psi,
          double2 *__restrict__ chi) {
  __shared__ ProjectLds<false> L;
  project_body<false>(g, vdt, ak, rur, rup, uz, psi, chi, nullptr, L);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_project_split(KGrid g, const float2 *__restrict__ vdt, const double *__restrict__ ak, const double2 *__restrict__ rur,
                const double2 *__restrict__ rup, const double2 *__restrict__ uz, double2 *__restrict__ psi,
                double2 *__restrict__ chi, const int *__restrict__ rowe) {
  __shared__ ProjectLds<true> L;
  project_body<true>(g, vdt, ak, rur, rup, uz, psi, chi, rowe, L);
}

int launch_project(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *psi, double2 *chi) {
  KGrid g = kgrid(c);
  const int ntc = (2 * g.nxchopdim + PJ_TC - 1) / PJ_TC, ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN;
  const double per_k = 5 * 2.0 * g.nr * g.nxchopdim;  // five products
  if (!c->leg_split) {
    KScope ks(c, K_PROJECT, per_k * (double)sum_nrchops(c));
    hipLaunchKernelGGL(k_project, dim3(xcd_grid(ntn * ntc * g.mcnt)), dim3(256), 0, c->stream, g, c->d.vdt, c->d.ak, rur,
                       rup, uz, psi, chi);
    MLG_CHECK_HIP(hipGetLastError());
    return 0;
  }
  KGrid g0, g1;
  split_modes(c, g, g0, g1);
  {
    KScope ks(c, K_PROJECT, per_k * split_fp64_sum(g0, 1));
    hipLaunchKernelGGL(k_project, dim3(xcd_grid(ntn * ntc * g0.mcnt)), dim3(256), 0, c->stream, g0, c->d.vdt, c->d.ak, rur,
                       rup, uz, psi, chi);
    MLG_CHECK_HIP(hipGetLastError());
  }
  if (g1.mcnt > 0) {
    const int ntn32 = (g.nrchopdim + 31) / 32;
    KScope ks(c, K_PROJECT, per_k * ((double)sum_nrchops(c) - split_fp64_sum(g0, 1)));
    hipLaunchKernelGGL(k_project_split, dim3(xcd_grid(ntn32 * ntc * g1.mcnt)), dim3(256), 0, c->stream, g1, c->d.vdt_s,
                       c->d.ak, rur, rup, uz, psi, chi, c->d.vdt_e);
    MLG_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

// -------------------------------------------------------------------------
// PROJECT of spectral velocities.  For FFF inputs, RTRAN backward followed by
// PROJECT's folded products is one linear map per m: with the folds
// F_e = 2 BE, F_o = 2 BO of the backward transform,
//   W[n] = sum_i T[n][i] F_p(i) = sum_{n' of parity p} C[n][n'] a[n'],
//   C[n][n'] = 2 sum_i T[n][i] PF(i, n'),
// for each table T in {V, D, T} (p = parity of n for V and T, the other
// parity for D).  The composed tables are built once (k_compose_vdt); the
// product kernel is PROJECT's with the spectral rows n' as its K extent
// (NRCHOPS(m) rows instead of NR/2 nodes), and the pieces that RTRAN and
// VEL2PLN add to the m = 0, k = 0 column only (log terms, VEL2PLN's
// correction) are added by k_project_fff_m0.
// Layout: cvdt[((t*NTCHOPDIM + m)*2 + p)*NH + j)*NH + h] holds C_t for the
// output row n = 2h + p and the input row n' = 2j + q(t, p), NH =
// ceil(NRCHOPDIM/2): 32 consecutive h of one (t, m, p, j) are one run.
// -------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_compose_vdt(KGrid g, const double *__restrict__ pf, const double *__restrict__ vdt,
                                                     double *__restrict__ cvdt) {
  extern __shared__ double prow[];  // PF row n' [NRH]
  const int nh = (g.nrchopdim + 1) / 2;
  const int j = blockIdx.x, pq = blockIdx.y, mm = blockIdx.z / 3, t = blockIdx.z % 3;
  const int p = pq;                            // parity of the output rows
  const int q = (t == 1) ? 1 - p : p;          // parity of the input rows
  const int np = 2 * j + q;
  const bool have = np < g.nrchopdim;
  if (have)
    for (int i = threadIdx.x; i < g.nrh; i += blockDim.x) prow[i] = pf[((size_t)mm * g.npf + np) * g.nrh + i];
  __syncthreads();
  const double *vt = vdt + ((size_t)t * g.ntchopdim + mm) * (size_t)g.nrchopdim * g.nrh;
  double *out = cvdt + ((((size_t)t * g.ntchopdim + mm) * 2 + p) * nh + j) * nh;
  for (int h = threadIdx.x; h < nh; h += blockDim.x) {
    const int n = 2 * h + p;
    double acc = 0.0;
    if (have && n < g.nrchopdim) {
      const double *vr = vt + (size_t)n * g.nrh;
      for (int i = 0; i < g.nrh; i++) acc += vr[i] * prow[i];
    }
    out[h] = 2.0 * acc;
  }
}

template <bool SPLIT>
__device__ __forceinline__ void project_fff_body(const KGrid &g, const typename LegElem<SPLIT>::T *__restrict__ cvdt,
                                                 const double *__restrict__ ak, const double2 *__restrict__ rur,
                                                 const double2 *__restrict__ rup, const double2 *__restrict__ uz,
                                                 double2 *__restrict__ psi, double2 *__restrict__ chi,
                                                 const int *__restrict__ rowe, ProjectLds<SPLIT> &L) {
  typedef typename LegElem<SPLIT>::T TE;
  constexpr int TN = SPLIT ? 32 : PJ_TN, NA = SPLIT ? 1 : 2, HB = TN / 2, NQT = 3 * TN / 16;
  const int nk = g.nxchopdim, nh = (g.nrchopdim + 1) / 2;
  const int ntc = (2 * nk + PJ_TC - 1) / PJ_TC, ntn = (g.nrchopdim + TN - 1) / TN;
  const int lid = xcd_logical(ntc * ntn * g.mcnt);
  if (lid < 0) return;
  const int ml = lid / (ntc * ntn), rem = lid - ml * (ntc * ntn), mm = g.mlo + ml;
  const int tn = rem / ntc, tc = rem - tn * ntc;
  const int mv = mm * g.minc;
  const int nn = mm < g.ntchop ? d_nrchops(g.nrchop, mv) : 0;
  const int n0 = tn * TN, c0 = tc * PJ_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n0 >= nn) {
    project_zero_tile<TN>(g, mm, n0, c0, psi, chi);
    return;
  }
  const int par = wave & 1, ch = wave >> 1;

  v4d acc[NA][5];
  v4f hh[NA][5], xx[NA][5];
  for (int a = 0; a < NA; a++)
    for (int q = 0; q < 5; q++) {
      acc[a][q] = (v4d){0, 0, 0, 0};
      hh[a][q] = xx[a][q] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
  const double2 *src[3] = {rur, rup, uz};
  TE rt[NQT];
  double2 rd[6];
  const int nchr = (nh + PJ_IC - 1) / PJ_IC;  // chunks of a split table row
  int rre = -1022;
  const int kstride = g.nrchopdim * g.ntchopdim;
  const size_t tstride = (size_t)g.ntchopdim * 2 * nh * nh;  // one table
  const TE *cm = cvdt + (size_t)mm * 2 * nh * nh;
  const int h0 = n0 >> 1;
  const int njp = (nn + 1) / 2;  // input row pairs
  auto fetch = [&](int j0) {
    if constexpr (SPLIT)
      if (tid < 3 * TN) {  // [table 3][h HB][parity 2]
        const int t = tid / TN, p = tid & 1, h = h0 + (tid % TN) / 2;
        rre = 2 * h + p < nn ? rowe[((((size_t)t * g.ntchopdim + mm) * 2 + p) * nh + h) * nchr + j0 / PJ_IC] : -1022;
      }
#pragma unroll
    for (int q = 0; q < NQT; q++) {
      const int idx = tid + 256 * q;  // [table 3][pair 16][parity 2][h HB]
      const int hl = idx % HB, p = (idx / HB) & 1, jl = (idx / (2 * HB)) & 15, t = idx / (32 * HB);
      const int n = n0 + 2 * hl + p, np = 2 * (j0 + jl) + ((t == 1) ? 1 - p : p);
      rt[q] = (n < nn && np < nn) ? cm[(size_t)t * tstride + ((size_t)p * nh + j0 + jl) * nh + h0 + hl] : TE{};
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int idx = tid + 256 * q;  // [field 3][k 16][row 32]
      const int rl = idx & 31, kl = (idx >> 5) & 15, f = idx >> 9;
      const int np = 2 * j0 + rl, k = (c0 >> 1) + kl;
      rd[q] = (np < nn && k < nk) ? src[f][(size_t)g.nrchopdim * mm + np + (size_t)kstride * k] : make_double2(0.0, 0.0);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < NQT; q++) {
      const int idx = tid + 256 * q;
      const int hl = idx % HB, p = (idx / HB) & 1, jl = (idx / (2 * HB)) & 15, t = idx / (32 * HB);
      L.k.T[t][p][jl][hl ^ lsw(jl, p)] = rt[q];
    }
    if constexpr (SPLIT)
      if (tid < 3 * TN) L.k.re[tid / TN][tid & 1][(tid % TN) / 2] = rre;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int idx = tid + 256 * q;
      const int rl = idx & 31, kl = (idx >> 5) & 15, f = idx >> 9;
      const int jl = rl >> 1, qp = rl & 1;
      const int sw = lsw(jl, 0);
      if constexpr (SPLIT) {
        // the chunk's 32 rows of (field, column) lie on 32 adjacent lanes
        const int e = grid_e(group_max<32>(max(bexp(rd[q].x), bexp(rd[q].y))));
        const double inv = inv_grid(e);
        if (rl == 0) L.k.ce[f][0][kl] = e;
        L.k.F[f][qp][jl][(2 * kl) ^ sw] = split_hl(rd[q].x, inv);
        L.k.F[f][qp][jl][(2 * kl + 1) ^ sw] = split_hl(rd[q].y, inv);
      } else {
        L.k.F[f][qp][jl][(2 * kl) ^ sw] = rd[q].x;
        L.k.F[f][qp][jl][(2 * kl + 1) ^ sw] = rd[q].y;
      }
    }
  };
  const int eo = par, oe = par ^ 1;
  // subtiles that hold data: rows n0 + 32 a + par + 2 l, columns c0 + 16 ch + l
  const bool pa0 = n0 + par < nn, pa1 = NA > 1 && n0 + 32 + par < nn, pc0 = c0 + ch * 16 < 2 * nk;
  fetch(0);
  for (int j0 = 0; j0 < njp; j0 += PJ_IC) {
    __syncthreads();
    stash();
    __syncthreads();
    if (j0 + PJ_IC < njp) fetch(j0 + PJ_IC);
    if (!(pa0 && pc0)) continue;  // all the wave's rows or columns are padding
    const int jleft = njp - j0;
#pragma unroll
    for (int ks = 0; ks < PJ_IC; ks += 4) {
      if (ks >= jleft) break;  // zero padding at the end of K
      const int kr = ks + (lane >> 4), cl = (ch * 16 + (lane & 15)) ^ lsw(kr, 0), st = lsw(kr, par);
      const TE f0e = L.k.F[0][eo][kr][cl], f0o = L.k.F[0][oe][kr][cl];
      const TE f1e = L.k.F[1][eo][kr][cl], f1o = L.k.F[1][oe][kr][cl];
      const TE f2e = L.k.F[2][eo][kr][cl];
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (a == 1 && !pa1) break;
        const int rl = (16 * a + (lane & 15)) ^ st;
        const TE av = L.k.T[0][par][kr][rl], ad = L.k.T[1][par][kr][rl], at = L.k.T[2][par][kr][rl];
        if constexpr (SPLIT) {
          mfma_split(av, f0e, hh[a][0], xx[a][0]);
          mfma_split(ad, f1o, hh[a][1], xx[a][1]);
          mfma_split(ad, f0o, hh[a][2], xx[a][2]);
          mfma_split(av, f1e, hh[a][3], xx[a][3]);
          mfma_split(at, f2e, hh[a][4], xx[a][4]);
        } else {
          acc[a][0] = mfma16x16x4(av, f0e, acc[a][0]);
          acc[a][1] = mfma16x16x4(ad, f1o, acc[a][1]);
          acc[a][2] = mfma16x16x4(ad, f0o, acc[a][2]);
          acc[a][3] = mfma16x16x4(av, f1e, acc[a][3]);
          acc[a][4] = mfma16x16x4(at, f2e, acc[a][4]);
        }
      }
    }
    if constexpr (SPLIT) {
      // one 16-row subtile per wave: rows h0 + 4 (lane / 16) + v of its parity
      const int kl = (ch * 16 + (lane & 15)) >> 1, rb = 4 * (lane >> 4);
      const int4 ev = *(const int4 *)&L.k.re[0][par][rb], ed = *(const int4 *)&L.k.re[1][par][rb];
      const int4 et = *(const int4 *)&L.k.re[2][par][rb];
      const int e0 = L.k.ce[0][0][kl], e1 = L.k.ce[1][0][kl];
      fold_split(acc[0][0], hh[0][0], xx[0][0], ev, e0);
      fold_split(acc[0][1], hh[0][1], xx[0][1], ed, e1);
      fold_split(acc[0][2], hh[0][2], xx[0][2], ed, e0);
      fold_split(acc[0][3], hh[0][3], xx[0][3], ev, e1);
      fold_split(acc[0][4], hh[0][4], xx[0][4], et, L.k.ce[2][0][kl]);
    }
  }
  __syncthreads();
  // epilogue as in PROJECT: psi = -i m W1 - W2, del^2 chi = i k W3 + m k W4 - W5
  project_epilogue<SPLIT, TN, NA>(g, L, acc, ak, mm, mv, nn, n0, c0, par, ch, psi, chi);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_project_fff(KGrid g, const double *__restrict__ cvdt, const double *__restrict__ ak, const double2 *__restrict__ rur,
              const double2 *__restrict__ rup, const double2 *__restrict__ uz, double2 *__restrict__ psi,
              double2 *__restrict__ chi) {
  __shared__ ProjectLds<false> L;
  project_fff_body<false>(g, cvdt, ak, rur, rup, uz, psi, chi, nullptr, L);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_project_fff_split(KGrid g, const float2 *__restrict__ cvdt, const double *__restrict__ ak,
                    const double2 *__restrict__ rur, const double2 *__restrict__ rup, const double2 *__restrict__ uz,
                    double2 *__restrict__ psi, double2 *__restrict__ chi, const int *__restrict__ rowe) {
  __shared__ ProjectLds<true> L;
  project_fff_body<true>(g, cvdt, ak, rur, rup, uz, psi, chi, rowe, L);
}

// The m = 0, k = 0 column's additions: RTRAN's log terms LN(u) (-log(1-x))
// of r u_theta and u_z through the D and T tables, and VEL2PLN: the log
// coefficient LN(psi) = -u(inf)/2 from the folds of Re(r u_theta) (with its
// log term), and the removal of its log image through the D table.  At
// m = 0, k = 0 psi = -W2 and del^2 chi = -W5.  All of it is linear in the
// column, so it is applied with vectors built from the tables (k_m0_vectors):
//   pln = -(sum_n LF[n] Re a[n] + LN(r u_theta) LFL) / 2,
//   LF[n] = 2 sum_i PF(i, n) G_p(n)(i),  G_e/o(i) = sum_{n' even/odd < NN}
//           exp(lognorm n') w_i PF(i, n'),  LFL = sum_i (lnx_e G_e + lnx_o G_o)(i),
//   psi[n] -= LN(r u_theta) LD[n] + pln DC[n],  del^2 chi[n] -= LN(u_z) LT[n],
// with lnx_e/o the folds of -log(1-x) and LD, LT, DC the D, T rows against
// the folds of -log(1-x) and of VEL2PLN's log image.
// Vector store (NRCHOPDIM each): LF, LD, LT, DC, then G_e, G_o (NRH), LFL.
__device__ __forceinline__ double block_sum256(double v, double *part) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((part[0] + part[1]) + part[2]) + part[3];
}
// stage 1 (one block per node i): G_e(i), G_o(i) over n' < nn
__global__ void __launch_bounds__(256) k_m0_gauss(KGrid g, int nn, const double *__restrict__ pf,
                                                  const double *__restrict__ w, const double *__restrict__ lognorm,
                                                  double *__restrict__ vec) {
  __shared__ double part[4];
  const int i = blockIdx.x, nrd = g.nrchopdim;
  double se = 0.0, so = 0.0;
  for (int n = threadIdx.x; n < nn; n += 256) {
    const double v = exp(lognorm[n]) * w[i] * pf[(size_t)n * g.nrh + i];
    if (n & 1) so += v;
    else se += v;
  }
  se = block_sum256(se, part);
  so = block_sum256(so, part);
  if (threadIdx.x == 0) {
    vec[4 * nrd + i] = se;
    vec[4 * nrd + g.nrh + i] = so;
  }
}
// stage 2 (one block per row n, and one more for LFL): LF, LD, LT, DC
__global__ void __launch_bounds__(256) k_m0_rows(KGrid g, const double *__restrict__ pf, const double *__restrict__ vdt,
                                                 const double *__restrict__ lnx, double *__restrict__ vec) {
  __shared__ double part[4];
  const int n = blockIdx.x, nrd = g.nrchopdim;
  const double *ge = vec + 4 * nrd, *go = ge + g.nrh;
  if (n == nrd) {  // LFL
    double s = 0.0;
    for (int i = threadIdx.x; i < g.nrh; i += 256) {
      const double la = lnx[i], lb = lnx[g.nr - 1 - i];
      s += (la + lb) * ge[i] + (la - lb) * go[i];
    }
    s = block_sum256(s, part);
    if (threadIdx.x == 0) vec[4 * nrd + 2 * g.nrh] = s;
    return;
  }
  const double *dr = vdt + ((size_t)1 * g.ntchopdim * nrd + n) * g.nrh;  // D, m = 0
  const double *tr = vdt + ((size_t)2 * g.ntchopdim * nrd + n) * g.nrh;  // T, m = 0
  const double *pn = pf + (size_t)n * g.nrh;
  const bool odd = n & 1;
  double lf = 0.0, ld = 0.0, lt = 0.0, dc = 0.0;
  for (int i = threadIdx.x; i < g.nrh; i += 256) {
    const double la = lnx[i], lb = lnx[g.nr - 1 - i];
    lf += pn[i] * (odd ? go[i] : ge[i]);
    // D pairs row n with the other parity's fold, T with its own
    ld += dr[i] * (odd ? (la + lb) : (la - lb));
    lt += tr[i] * (odd ? (la - lb) : (la + lb));
    dc += dr[i] * (odd ? pf[i] / g.norm0 : pf[(size_t)g.nrh + i] / g.norm1);
  }
  lf = block_sum256(lf, part);
  ld = block_sum256(ld, part);
  lt = block_sum256(lt, part);
  dc = block_sum256(dc, part);
  if (threadIdx.x == 0) {
    vec[n] = 2.0 * lf;
    vec[nrd + n] = ld;
    vec[2 * nrd + n] = lt;
    vec[3 * nrd + n] = 2.0 * dc;
  }
}
// the additions themselves (one block)
// Block k (grid: NXCHOPDIM blocks) then applies PROJECT's value-at-infinity
// fix-up to psi's (m = 0, k) column, after block 0's additions to (0, 0):
// A(0,0,k) -= sum_n AT1(n) A(n,0,k) / AT1(0) (:2541-2543), the products
// formed in parallel and summed by one thread in ascending n.
__global__ void __launch_bounds__(256) k_project_fff_m0(KGrid g, const double *__restrict__ vec, const double2 *__restrict__ rup,
                                                        const double *__restrict__ ln_rup, const double *__restrict__ ln_uz,
                                                        double2 *__restrict__ psi, double *__restrict__ psi_ln,
                                                        double2 *__restrict__ chi, int ln_mode, double ln,
                                                        const double *__restrict__ at1) {
  __shared__ double part[4];
  extern __shared__ __align__(16) double2 fsp[];  // [NRCHOP] products of the fix-up
  if (blockIdx.x == 0) {
    const int nn = g.ntchop >= 1 ? g.nrchop : 0, nrd = g.nrchopdim;
    const double lr = ln_rup[0], lz = ln_uz[0];
    double s = 0.0;
    for (int n = threadIdx.x; n < nn; n += 256) s += vec[n] * rup[n].x;
    s = block_sum256(s, part);
    const double pln = ln_mode != 0 ? -(s + lr * vec[4 * nrd + 2 * g.nrh]) / 2.0 : 0.0;  // VEL2PLN (ln_mode 0: none)
    for (int n = threadIdx.x; n < nn; n += 256) {
      psi[n].x = psi[n].x - (lr * vec[nrd + n] + pln * vec[3 * nrd + n]);
      chi[n].x = chi[n].x - lz * vec[2 * nrd + n];
    }
    if (threadIdx.x == 0 && ln_mode != 0 && ln_mode != 2) psi_ln[0] = ln_mode == 1 ? ln : pln;
    __syncthreads();
  }
  if (g.nrchop < 1) return;
  double2 *col = psi + (size_t)g.nrchopdim * (size_t)g.ntchopdim * blockIdx.x;
  for (int n = threadIdx.x; n < g.nrchop; n += 256) fsp[n] = make_double2(at1[n] * col[n].x, at1[n] * col[n].y);
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0.0, sy = 0.0;
#pragma unroll 8
    for (int n = 0; n < g.nrchop; n++) {
      sx = sx + fsp[n].x;
      sy = sy + fsp[n].y;
    }
    const double2 v0 = col[0];
    col[0] = make_double2(v0.x - sx / g.at1_0, v0.y - sy / g.at1_0);
  }
}

// The composed tables (once) and the m = 0 vectors (once per truncation: a
// buffer per NRCHOP value, kept until the context ends, so that a captured
// step that reads one set is never handed another truncation's vectors).
static int ensure_cvdt(mlg_ctx *c, const double **m0vec) {
  if (c->g.nrh < 1) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (!c->d.cvdt) {
    const int nh = (g.nrchopdim + 1) / 2;
    const size_t n = (size_t)3 * g.ntchopdim * 2 * nh * nh;
    MLG_CHECK_HIP(hipMalloc((void **)&c->d.cvdt, sizeof(double) * n));
    hipLaunchKernelGGL(k_compose_vdt, dim3(nh, 2, 3 * g.ntchopdim), dim3(256), sizeof(double) * (size_t)g.nrh, c->stream, g,
                       c->d.pf, c->d.vdt, c->d.cvdt);
    MLG_CHECK_HIP(hipGetLastError());
  }
  if (c->leg_split && !c->d.cvdt_s) {
    // composed rows (t, m, p, h) run over j with stride NH
    const int nh = (g.nrchopdim + 1) / 2;
    const long n = (long)3 * g.ntchopdim * 2 * nh * nh;
    const int rc = split_rows(c, c->d.cvdt, n, &c->d.cvdt_s, &c->d.cvdt_e, (long)3 * g.ntchopdim * 2 * nh, nh,
                              (long)nh * nh, 1, nh, nh, SPLIT_KC);
    if (rc) return rc;
  }
  for (auto &v : c->m0vecs)
    if (v.first == g.nrchop) {
      *m0vec = v.second;
      return 0;
    }
  double *m0 = nullptr;
  MLG_CHECK_HIP(hipMalloc((void **)&m0, sizeof(double) * (4 * (size_t)g.nrchopdim + 2 * (size_t)g.nrh + 1)));
  c->m0vecs.emplace_back(g.nrchop, m0);
  const int nn = g.ntchop >= 1 ? g.nrchop : 0;
  hipLaunchKernelGGL(k_m0_gauss, dim3(g.nrh), dim3(256), 0, c->stream, g, nn, c->d.pf, c->d.w, c->d.lognorm, m0);
  hipLaunchKernelGGL(k_m0_rows, dim3(g.nrchopdim + 1), dim3(256), 0, c->stream, g, c->d.pf, c->d.vdt, c->d.lnx, m0);
  MLG_CHECK_HIP(hipGetLastError());
  *m0vec = m0;
  return 0;
}

int launch_project_fff(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, const double *const ln_u[3],
                       double2 *psi, double *psi_ln, double2 *chi, int ln_mode, double ln) {
  const double *m0vec = nullptr;
  const int rc = ensure_cvdt(c, &m0vec);
  if (rc) return rc;
  KGrid g = kgrid(c);
  const int ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN, ntc = (2 * g.nxchopdim + PJ_TC - 1) / PJ_TC;
  // five products of NRCHOPS(m) rows x NRCHOPS(m)/2 same-class input rows per real column
  const double per_k = 5 * 2.0 * g.nxchopdim;
  if (!c->leg_split) {
    KScope ks(c, K_PROJECT, per_k * sum_nrchops_sq(c));
    hipLaunchKernelGGL(k_project_fff, dim3(xcd_grid(ntn * ntc * g.mcnt)), dim3(256), 0, c->stream, g, c->d.cvdt, c->d.ak,
                       rur, rup, uz, psi, chi);
    MLG_CHECK_HIP(hipGetLastError());
  } else {
    KGrid g0, g1;
    split_modes(c, g, g0, g1);
    const double sq0 = split_fp64_sum(g0, 2);
    {
      KScope ks(c, K_PROJECT, per_k * sq0);
      hipLaunchKernelGGL(k_project_fff, dim3(xcd_grid(ntn * ntc * g0.mcnt)), dim3(256), 0, c->stream, g0, c->d.cvdt, c->d.ak,
                         rur, rup, uz, psi, chi);
      MLG_CHECK_HIP(hipGetLastError());
    }
    if (g1.mcnt > 0) {
      const int ntn32 = (g.nrchopdim + 31) / 32;
      KScope ks(c, K_PROJECT, per_k * (sum_nrchops_sq(c) - sq0));
      hipLaunchKernelGGL(k_project_fff_split, dim3(xcd_grid(ntn32 * ntc * g1.mcnt)), dim3(256), 0, c->stream, g1,
                         c->d.cvdt_s, c->d.ak, rur, rup, uz, psi, chi, c->d.cvdt_e);
      MLG_CHECK_HIP(hipGetLastError());
    }
  }
  // with the value-at-infinity fix-up of psi (one block per k)
  hipLaunchKernelGGL(k_project_fff_m0, dim3(g.nxchopdim), dim3(256), sizeof(double2) * (size_t)(g.nrchop > 0 ? g.nrchop : 1),
                     c->stream, g, m0vec, rup, ln_u[1], ln_u[2], psi, psi_ln, chi, ln_mode, ln, c->d.at1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
