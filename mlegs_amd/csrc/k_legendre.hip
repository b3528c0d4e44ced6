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
// Backward: FFF (n,m,k) -> PFF (r,m,k).
// BE(i,c) = sum_j PF(i,2j) a(2j,c), BO(i,c) = sum_j PF(i,2j+1) a(2j+1,c);
// f(x_i) = BE+BO, f(x_{NR-1-i}) = BE-BO; LN*(-log(1-x)) added at m=0,k=0.
// Block tile: 64 radial nodes x 64 real columns (32 complex k); 4 waves as
// 2x2, each wave 32x32 = 2x2 MFMA tiles per parity.  K chunk: 16 pairs.
// -------------------------------------------------------------------------
#define BW_TI 64
#define BW_TC 64
#define BW_JC 16

// LDS: one K-chunk stage of the table and data tiles (the next chunk waits
// in registers); after the K loop the same memory stages the output, one
// half (S = BE+BO, then D = BE-BO) at a time.  33 KB per block: four blocks
// per CU.
union RtranBwdLds {
  struct {
    double A[2][BW_JC][BW_TI];  // [parity][j][i]
    double B[2][BW_JC][BW_TC];  // [parity][j][col]
  } k;
  double E[BW_TC][BW_TI + 1];  // [col][i]
};

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_rtran_bwd(KGrid g, const double *__restrict__ pf, TransformJobs jobs, int nj, const double *__restrict__ lnx) {
  // logical block (radial tile fastest, then field, column tile, m): the
  // radial tiles that read one data tile run one after another on one XCD
  // (the data come from L2 after the first), and the mode's table tiles
  // (all radial tiles of m) stay in that L2 for every field and column tile
  const int nk = g.nxchopdim;
  const int ntiles_c = (2 * nk + BW_TC - 1) / BW_TC, ntiles_i = (g.nrh + BW_TI - 1) / BW_TI;
  const int lid = xcd_logical(ntiles_c * ntiles_i * g.ntchopdim * nj);
  if (lid < 0) return;
  const int ti = lid % ntiles_i, r1 = lid / ntiles_i;
  const int fj = r1 % nj, r2 = r1 / nj;
  const int tc = r2 % ntiles_c, mm = r2 / ntiles_c;
  const double2 *__restrict__ in = jobs.in[fj];
  const double *__restrict__ ln_in = jobs.ln[fj];
  double2 *__restrict__ out = jobs.out[fj];
  const int mv = mm * g.minc;
  const int nn = d_nrchops(g.nrchop, mv);
  const int i0 = ti * BW_TI, c0 = tc * BW_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave & 1, wc = wave >> 1;
  __shared__ RtranBwdLds L;

  v4d acc[2][2][2];  // [parity][sub i][sub c]
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) acc[p][a][b] = (v4d){0, 0, 0, 0};

  const double *pfm = pf + (size_t)mm * g.npf * g.nrh;
  const int njp = (nn + 1) / 2;  // (even, odd) pairs
  const int nch = (njp + BW_JC - 1) / BW_JC;
  // per-thread prefetch registers: 8 table values, 4 complex data values
  double ra[8];
  double2 rb[4];
  // 32-bit element offsets: lane part fixed, chunk and q parts uniform
  const int ta_i = i0 + (tid & 63), ta_n = tid >> 6;          // table: i, n - 2 j0 - 4 q
  const int tb_n = tid & 31, tb_k = (c0 >> 1) + (tid >> 5);   // data: n - 2 j0, k - 8 q
  const int kstride = g.nrchopdim * g.ntchopdim;
  const double2 *inm = in + (size_t)g.nrchopdim * mm;
  auto fetch = [&](int ch) {
    const int j0 = ch * BW_JC;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int n = 2 * j0 + ta_n + 4 * q;
      ra[q] = (n < nn && ta_i < g.nrh) ? pfm[n * g.nrh + ta_i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = 2 * j0 + tb_n, k = tb_k + 8 * q;
      rb[q] = (n < nn && k < nk) ? inm[n + kstride * k] : make_double2(0.0, 0.0);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int idx = tid + 256 * q;
      int i = idx & 63, nl = idx >> 6;
      L.k.A[nl & 1][nl >> 1][i ^ lsw(nl >> 1, 0)] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int idx = tid + 256 * q;
      int nl = idx & 31, kl = idx >> 5;
      const int sw = lsw(nl >> 1, nl & 1);
      L.k.B[nl & 1][nl >> 1][(2 * kl) ^ sw] = rb[q].x;
      L.k.B[nl & 1][nl >> 1][(2 * kl + 1) ^ sw] = rb[q].y;
    }
  };
  // the wave's column subtiles that hold data (the last tile of a mode is
  // partly padding); the others skip their products (wave-uniform)
  const bool live0 = c0 + wc * 32 < 2 * nk, live1 = c0 + wc * 32 + 16 < 2 * nk;
  if (nch > 0) fetch(0);
  for (int ch = 0; ch < nch; ch++) {
    __syncthreads();
    stash();
    __syncthreads();
    if (ch + 1 < nch) fetch(ch + 1);
    const int jleft = njp - ch * BW_JC;  // pairs of this chunk that hold data
#pragma unroll
    for (int ks = 0; ks < BW_JC; ks += 4) {
      if (ks >= jleft || !live0) break;  // zero padding (chunk tail, or all the wave's columns)
      const int kr = ks + (lane >> 4);
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const int sa = lsw(kr, 0), sb = lsw(kr, p);
        double a0 = L.k.A[p][kr][(wi * 32 + (lane & 15)) ^ sa];
        double a1 = L.k.A[p][kr][(wi * 32 + 16 + (lane & 15)) ^ sa];
        double b0 = L.k.B[p][kr][(wc * 32 + (lane & 15)) ^ sb];
        double b1 = L.k.B[p][kr][(wc * 32 + 16 + (lane & 15)) ^ sb];
        acc[p][0][0] = mfma16x16x4(a0, b0, acc[p][0][0]);
        acc[p][1][0] = mfma16x16x4(a1, b0, acc[p][1][0]);
        if (live1) {
          acc[p][0][1] = mfma16x16x4(a0, b1, acc[p][0][1]);
          acc[p][1][1] = mfma16x16x4(a1, b1, acc[p][1][1]);
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
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          int il = wi * 32 + a * 16 + (lane >> 4) + 4 * v;
          int cl = wc * 32 + b * 16 + (lane & 15);
          double e = acc[0][a][b][v], o = acc[1][a][b][v];
          L.E[cl][il] = h == 0 ? e + o : e - o;
        }
    __syncthreads();
    for (int q = 0; q < 8; q++) {
      int idx = tid + 256 * q;  // 0..2047: i fastest
      int il = idx & 63, kl = idx >> 6;
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
    for (int idx = tid; idx < pad * (BW_TC / 2); idx += 256) {
      const int kl = idx / pad, k = (c0 >> 1) + kl;
      if (k < nk) out[(size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * k) + g.nr + (idx - kl * pad)] = make_double2(0.0, 0.0);
    }
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
  int nti = (g.nrh + BW_TI - 1) / BW_TI, ntc = (2 * g.nxchopdim + BW_TC - 1) / BW_TC;
  dim3 grid(xcd_grid(nti * ntc * g.ntchopdim * n));
  // 2 NR sum_m NRCHOPS(m) NXCHOPDIM flops per field (real table x complex data, folded)
  KScope ks(c, K_RTRAN_BWD, n * 2.0 * g.nr * (double)sum_nrchops(c) * g.nxchopdim);
  hipLaunchKernelGGL(k_rtran_bwd, grid, dim3(256), 0, c->stream, g, c->d.pf, jobs, n, c->d.lnx);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_rtran_bwd(mlg_ctx *c, const double2 *in, const double *ln_in, double2 *out) {
  return launch_rtran_bwd_n(c, 1, &in, &ln_in, &out);
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

union RtranFwdLds {
  struct {
    double A[2][FW_IC][FW_TN / 2];  // [parity][i][n-local within parity]
    double B[2][FW_IC][FW_TC];      // [fold parity][i][col]
  } k;
  double C[FW_TC][FW_TN + 1];  // [col][n-local]
};

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_rtran_fwd(KGrid g, const double *__restrict__ pf, const double *__restrict__ w, const double2 *__restrict__ in,
            double2 *__restrict__ out, double *zl) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && zl) zl[0] = 0.0;  // the result's LN
  const int nk = g.nxchopdim;
  const int ntc = (2 * nk + FW_TC - 1) / FW_TC, ntn = (g.nrchopdim + FW_TN - 1) / FW_TN;
  const int lid = xcd_logical(ntc * ntn * g.ntchopdim);
  if (lid < 0) return;
  const int mm = lid / (ntc * ntn), rem = lid - mm * (ntc * ntn);
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
  __shared__ RtranFwdLds L;

  v4d acc[2][2] = {{(v4d){0, 0, 0, 0}, (v4d){0, 0, 0, 0}}, {(v4d){0, 0, 0, 0}, (v4d){0, 0, 0, 0}}};
  const double *pfm = pf + (size_t)mm * g.npf * g.nrh;
  const double2 *inm = in + (size_t)g.ndimr * mm;
  const int kstride = g.ndimr * g.ntchopdim;
  // lane parts of the fetch indices (table: i fastest; data: i fastest)
  const int il = tid & 15, nl0 = tid >> 4, kl0 = tid >> 4;
  double ra[4];
  double2 rp[2], rq[2];
  double rw = 0.0;
  auto fetch = [&](int ib) {
    const int i = ib + il;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = n0 + nl0 + 16 * q;
      ra[q] = (n < nn && i < g.nrh) ? pfm[n * g.nrh + i] : 0.0;
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
    const int sw = lsw(il, 0);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int kl = kl0 + 16 * q;
      const double2 p = rp[q], r = rq[q];
      L.k.B[0][il][(2 * kl) ^ sw] = (p.x + r.x) * rw;
      L.k.B[0][il][(2 * kl + 1) ^ sw] = (p.y + r.y) * rw;
      L.k.B[1][il][(2 * kl) ^ sw] = (p.x - r.x) * rw;
      L.k.B[1][il][(2 * kl + 1) ^ sw] = (p.y - r.y) * rw;
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
      const double a0 = L.k.A[par][kr][(lane & 15) ^ swa];
      const double a1 = L.k.A[par][kr][(16 + (lane & 15)) ^ swa];
      const double b0 = L.k.B[par][kr][(ch * 32 + (lane & 15)) ^ sw];
      const double b1 = L.k.B[par][kr][(ch * 32 + 16 + (lane & 15)) ^ sw];
      acc[0][0] = mfma16x16x4(a0, b0, acc[0][0]);
      if (rc1) acc[0][1] = mfma16x16x4(a0, b1, acc[0][1]);
      if (ra1) {
        acc[1][0] = mfma16x16x4(a1, b0, acc[1][0]);
        if (rc1) acc[1][1] = mfma16x16x4(a1, b1, acc[1][1]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int rl = 16 * a + (lane >> 4) + 4 * v;  // row within the parity tile
        const int nl = 2 * rl + par;                    // n-local
        const int cl = ch * 32 + b * 16 + (lane & 15);
        L.C[cl][nl] = acc[a][b][v];
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

int launch_rtran_fwd(mlg_ctx *c, const double2 *in, double2 *out, double *zl) {
  KGrid g = kgrid(c);
  int ntn = (g.nrchopdim + FW_TN - 1) / FW_TN, ntc = (2 * g.nxchopdim + FW_TC - 1) / FW_TC;
  dim3 grid(xcd_grid(ntn * ntc * g.ntchopdim));
  KScope ks(c, K_RTRAN_FWD, 2.0 * g.nr * (double)sum_nrchops(c) * g.nxchopdim);
  hipLaunchKernelGGL(k_rtran_fwd, grid, dim3(256), 0, c->stream, g, c->d.pf, c->d.w, in, out, zl);
  MLG_CHECK_HIP(hipGetLastError());
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

union ProjectLds {
  struct {
    double T[3][2][PJ_IC][PJ_TN / 2];  // [V,D,T][parity][i][n-local within parity]
    double F[3][2][PJ_IC][PJ_TC];      // [field][fold parity][i][col]
  } k;
  struct {
    double P[PJ_TC / 2][PJ_TN + 1][2];  // psi   [complex k][n][re,im]
    double C[PJ_TC / 2][PJ_TN + 1][2];  // del^2 chi
  } e;
};

// Block tile: 64 rows n (32 even + 32 odd) x 32 real columns (16 complex k);
// wave w: parity w&1, column half w>>1, two 16-row subtiles.  The K loop
// (Gauss nodes, chunks of 16) stages one chunk in LDS while the next one is
// being fetched into registers.  The epilogue forms psi and del^2 chi in
// registers, exchanging re/im partners with a lane shuffle.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) k_project(KGrid g, const double *__restrict__ vdt, const double *__restrict__ ak,
                                                 const double2 *__restrict__ rur, const double2 *__restrict__ rup,
                                                 const double2 *__restrict__ uz, double2 *__restrict__ psi,
                                                 double2 *__restrict__ chi) {
  const int nk = g.nxchopdim;
  const int ntc = (2 * nk + PJ_TC - 1) / PJ_TC, ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN;
  const int lid = xcd_logical(ntc * ntn * g.ntchopdim);
  if (lid < 0) return;
  const int mm = lid / (ntc * ntn), rem = lid - mm * (ntc * ntn);
  const int tn = rem / ntc, tc = rem - tn * ntc;
  const int mv = mm * g.minc;
  const int nn = mm < g.ntchop ? d_nrchops(g.nrchop, mv) : 0;
  const int n0 = tn * PJ_TN, c0 = tc * PJ_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n0 >= nn) {  // a tile of the truncated rows: zero (CHOPDO)
    for (int q = 0; q < 4; q++) {
      const int idx = tid + 256 * q, n = n0 + (idx & 63), k = (c0 >> 1) + (idx >> 6);
      if (n < g.nrchopdim && k < nk) {
        const size_t o = (size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
        psi[o] = make_double2(0.0, 0.0);
        chi[o] = make_double2(0.0, 0.0);
      }
    }
    return;
  }
  const int par = wave & 1, ch = wave >> 1;
  const size_t tstride = (size_t)g.ntchopdim * g.nrchopdim * g.nrh;
  __shared__ ProjectLds L;

  v4d acc[2][5];
  for (int a = 0; a < 2; a++)
    for (int q = 0; q < 5; q++) acc[a][q] = (v4d){0, 0, 0, 0};
  const double2 *src[3] = {rur, rup, uz};
  double rt[12];
  double2 rp[3], rq[3];
  // 32-bit element offsets from per-m bases
  const double *vdm = vdt + (size_t)mm * g.nrchopdim * g.nrh;
  const int its = (int)tstride, kstride = g.ndimr * g.ntchopdim;
  const size_t moff = (size_t)g.ndimr * mm;
  auto fetch = [&](int ib) {
#pragma unroll
    for (int q = 0; q < 12; q++) {
      int idx = tid + 256 * q;  // [table 3][n 64][i 16]
      int il = idx & 15, nl = (idx >> 4) & 63, t = idx >> 10;
      int n = n0 + nl, i = ib + il;
      rt[q] = (n < nn && i < g.nrh) ? vdm[t * its + n * g.nrh + i] : 0.0;
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
    for (int q = 0; q < 12; q++) {
      int idx = tid + 256 * q;
      int il = idx & 15, nl = (idx >> 4) & 63, t = idx >> 10;
      L.k.T[t][nl & 1][il][(nl >> 1) ^ lsw(il, nl & 1)] = rt[q];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int idx = tid + 256 * q;
      int il = idx & 15, kl = (idx >> 4) & 15, f = idx >> 8;
      double2 p = rp[q], r = rq[q];
      const int sw = lsw(il, 0);
      L.k.F[f][0][il][(2 * kl) ^ sw] = p.x + r.x;
      L.k.F[f][0][il][(2 * kl + 1) ^ sw] = p.y + r.y;
      L.k.F[f][1][il][(2 * kl) ^ sw] = p.x - r.x;
      L.k.F[f][1][il][(2 * kl + 1) ^ sw] = p.y - r.y;
    }
  };
  const int eo = par, oe = par ^ 1;
  // subtiles that hold data: rows n0 + 32 a + par + 2 l, columns c0 + 16 ch + l
  const bool pa0 = n0 + par < nn, pa1 = n0 + 32 + par < nn, pc0 = c0 + ch * 16 < 2 * nk;
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
      const double f0e = L.k.F[0][eo][kr][cl], f0o = L.k.F[0][oe][kr][cl];
      const double f1e = L.k.F[1][eo][kr][cl], f1o = L.k.F[1][oe][kr][cl];
      const double f2e = L.k.F[2][eo][kr][cl];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        if (a == 1 && !pa1) break;
        const int rl = (16 * a + (lane & 15)) ^ st;
        const double av = L.k.T[0][par][kr][rl], ad = L.k.T[1][par][kr][rl], at = L.k.T[2][par][kr][rl];
        acc[a][0] = mfma16x16x4(av, f0e, acc[a][0]);
        acc[a][1] = mfma16x16x4(ad, f1o, acc[a][1]);
        acc[a][2] = mfma16x16x4(ad, f0o, acc[a][2]);
        acc[a][3] = mfma16x16x4(av, f1e, acc[a][3]);
        acc[a][4] = mfma16x16x4(at, f2e, acc[a][4]);
      }
    }
  }
  __syncthreads();
  // epilogue: lane (row, col) holds W1..W5 at one real column; its re/im
  // partner is lane ^ 1.
  const int cl = ch * 16 + (lane & 15), kl = cl >> 1, isim = cl & 1;
  const int kg = (c0 >> 1) + kl;
  const double m = (double)mv;
  const double kv = (kg < nk) ? ak[mm + (size_t)g.ntchopdim * kg] : 0.0, mk = m * kv;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rl = (lane >> 4) + 4 * v + 16 * a;
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
  for (int q = 0; q < 4; q++) {
    int idx = tid + 256 * q;  // 0..1023: n fastest, 16 complex k
    int nl = idx & 63, kq = idx >> 6;
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

int launch_project(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *psi, double2 *chi) {
  KGrid g = kgrid(c);
  int ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN, ntc = (2 * g.nxchopdim + PJ_TC - 1) / PJ_TC;
  dim3 grid(xcd_grid(ntn * ntc * g.ntchopdim));
  KScope ks(c, K_PROJECT, 5 * 2.0 * g.nr * (double)sum_nrchops(c) * g.nxchopdim);  // five products
  hipLaunchKernelGGL(k_project, grid, dim3(256), 0, c->stream, g, c->d.vdt, c->d.ak, rur, rup, uz, psi, chi);
  MLG_CHECK_HIP(hipGetLastError());
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

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_project_fff(KGrid g, const double *__restrict__ cvdt, const double *__restrict__ ak, const double2 *__restrict__ rur,
              const double2 *__restrict__ rup, const double2 *__restrict__ uz, double2 *__restrict__ psi,
              double2 *__restrict__ chi) {
  const int nk = g.nxchopdim, nh = (g.nrchopdim + 1) / 2;
  const int ntc = (2 * nk + PJ_TC - 1) / PJ_TC, ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN;
  const int lid = xcd_logical(ntc * ntn * g.ntchopdim);
  if (lid < 0) return;
  const int mm = lid / (ntc * ntn), rem = lid - mm * (ntc * ntn);
  const int tn = rem / ntc, tc = rem - tn * ntc;
  const int mv = mm * g.minc;
  const int nn = mm < g.ntchop ? d_nrchops(g.nrchop, mv) : 0;
  const int n0 = tn * PJ_TN, c0 = tc * PJ_TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n0 >= nn) {  // a tile of the truncated rows: zero (CHOPDO)
    for (int q = 0; q < 4; q++) {
      const int idx = tid + 256 * q, n = n0 + (idx & 63), k = (c0 >> 1) + (idx >> 6);
      if (n < g.nrchopdim && k < nk) {
        const size_t o = (size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
        psi[o] = make_double2(0.0, 0.0);
        chi[o] = make_double2(0.0, 0.0);
      }
    }
    return;
  }
  const int par = wave & 1, ch = wave >> 1;
  __shared__ ProjectLds L;

  v4d acc[2][5];
  for (int a = 0; a < 2; a++)
    for (int q = 0; q < 5; q++) acc[a][q] = (v4d){0, 0, 0, 0};
  const double2 *src[3] = {rur, rup, uz};
  double rt[12];
  double2 rd[6];
  const int kstride = g.nrchopdim * g.ntchopdim;
  const size_t tstride = (size_t)g.ntchopdim * 2 * nh * nh;  // one table
  const double *cm = cvdt + (size_t)mm * 2 * nh * nh;
  const int h0 = n0 >> 1;
  const int njp = (nn + 1) / 2;  // input row pairs
  auto fetch = [&](int j0) {
#pragma unroll
    for (int q = 0; q < 12; q++) {
      const int idx = tid + 256 * q;  // [table 3][pair 16][parity 2][h 32]
      const int hl = idx & 31, p = (idx >> 5) & 1, jl = (idx >> 6) & 15, t = idx >> 10;
      const int n = n0 + 2 * hl + p, np = 2 * (j0 + jl) + ((t == 1) ? 1 - p : p);
      rt[q] = (n < nn && np < nn) ? cm[(size_t)t * tstride + ((size_t)p * nh + j0 + jl) * nh + h0 + hl] : 0.0;
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
    for (int q = 0; q < 12; q++) {
      const int idx = tid + 256 * q;
      const int hl = idx & 31, p = (idx >> 5) & 1, jl = (idx >> 6) & 15, t = idx >> 10;
      L.k.T[t][p][jl][hl ^ lsw(jl, p)] = rt[q];
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int idx = tid + 256 * q;
      const int rl = idx & 31, kl = (idx >> 5) & 15, f = idx >> 9;
      const int jl = rl >> 1, qp = rl & 1;
      const int sw = lsw(jl, 0);
      L.k.F[f][qp][jl][(2 * kl) ^ sw] = rd[q].x;
      L.k.F[f][qp][jl][(2 * kl + 1) ^ sw] = rd[q].y;
    }
  };
  const int eo = par, oe = par ^ 1;
  // subtiles that hold data: rows n0 + 32 a + par + 2 l, columns c0 + 16 ch + l
  const bool pa0 = n0 + par < nn, pa1 = n0 + 32 + par < nn, pc0 = c0 + ch * 16 < 2 * nk;
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
      const double f0e = L.k.F[0][eo][kr][cl], f0o = L.k.F[0][oe][kr][cl];
      const double f1e = L.k.F[1][eo][kr][cl], f1o = L.k.F[1][oe][kr][cl];
      const double f2e = L.k.F[2][eo][kr][cl];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        if (a == 1 && !pa1) break;
        const int rl = (16 * a + (lane & 15)) ^ st;
        const double av = L.k.T[0][par][kr][rl], ad = L.k.T[1][par][kr][rl], at = L.k.T[2][par][kr][rl];
        acc[a][0] = mfma16x16x4(av, f0e, acc[a][0]);
        acc[a][1] = mfma16x16x4(ad, f1o, acc[a][1]);
        acc[a][2] = mfma16x16x4(ad, f0o, acc[a][2]);
        acc[a][3] = mfma16x16x4(av, f1e, acc[a][3]);
        acc[a][4] = mfma16x16x4(at, f2e, acc[a][4]);
      }
    }
  }
  __syncthreads();
  // epilogue as in k_project: psi = -i m W1 - W2, del^2 chi = i k W3 + m k W4 - W5
  const int cl = ch * 16 + (lane & 15), kl = cl >> 1, isim = cl & 1;
  const int kg = (c0 >> 1) + kl;
  const double m = (double)mv;
  const double kv = (kg < nk) ? ak[mm + (size_t)g.ntchopdim * kg] : 0.0, mk = m * kv;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rl = (lane >> 4) + 4 * v + 16 * a;
      const int nl = 2 * rl + par;
      const double w1 = acc[a][0][v], w2 = acc[a][1][v], w3 = acc[a][2][v], w4 = acc[a][3][v], w5 = acc[a][4][v];
      const double w1p = __shfl_xor(w1, 1, 64), w3p = __shfl_xor(w3, 1, 64);
      double ps, cs;
      if (!isim) {
        ps = m * w1p - w2;
        cs = (-kv * w3p + mk * w4) - w5;
      } else {
        ps = -m * w1p - w2;
        cs = (kv * w3p + mk * w4) - w5;
      }
      L.e.P[kl][nl][isim] = ps;
      L.e.C[kl][nl][isim] = cs;
    }
  __syncthreads();
  for (int q = 0; q < 4; q++) {
    int idx = tid + 256 * q;  // 0..1023: n fastest, 16 complex k
    int nl = idx & 63, kq = idx >> 6;
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
  int ntn = (g.nrchopdim + PJ_TN - 1) / PJ_TN, ntc = (2 * g.nxchopdim + PJ_TC - 1) / PJ_TC;
  dim3 grid(xcd_grid(ntn * ntc * g.ntchopdim));
  {
    // five products of NRCHOPS(m) rows x NRCHOPS(m)/2 same-class input rows per real column
    KScope ks(c, K_PROJECT, 5 * sum_nrchops_sq(c) * 2.0 * g.nxchopdim);
    hipLaunchKernelGGL(k_project_fff, grid, dim3(256), 0, c->stream, g, c->d.cvdt, c->d.ak, rur, rup, uz, psi, chi);
    MLG_CHECK_HIP(hipGetLastError());
  }
  // with the value-at-infinity fix-up of psi (one block per k)
  hipLaunchKernelGGL(k_project_fff_m0, dim3(g.nxchopdim), dim3(256), sizeof(double2) * (size_t)(g.nrchop > 0 ? g.nrchop : 1),
                     c->stream, g, m0vec, rup, ln_u[1], ln_u[2], psi, psi_ln, chi, ln_mode, ln, c->d.at1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
