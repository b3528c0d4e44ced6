// Spectral-space (FFF) kernels: banded operators, banded solves, pointwise
// updates, filters, checks and reductions.
//
// FFF element (n, m, k) lives at n + NRCHOPDIM*(m + NTCHOPDIM*k).  Banded
// operators (src/mod_legops.f90) apply, per (m, k), a tri/penta-diagonal
// matrix truncated to NRCHOPS(m) rows, exactly as BANMULC does
// (src/mod_bandmat.f90:1199-1243): diagonal term first, then the off-diagonal
// terms in band-column order, entries outside the truncated matrix skipped.
// Composite operators (PC2VEL, PC2VOR, GRAD) are fused into one pass that
// recomputes the intermediate bands per output row.
#include "mlegs_internal.h"
#include "xcd.h"

namespace mlg {

__device__ __forceinline__ int s_nrchops(int nrchop, int m) {
  int v = nrchop - m < nrchop ? nrchop - m : nrchop;
  return v > 0 ? v : 0;
}
__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 axc(double a, double2 x) { return make_double2(a * x.x, a * x.y); }
__device__ __forceinline__ double2 add2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 sub2(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 ix(double s, double2 x) {  // (0 + i s) * x
  return make_double2(0.0 * x.x - s * x.y, 0.0 * x.y + s * x.x);
}

struct FIdx {
  int n, mm, k;
};
// 32-bit decode (every FFF field has fewer than 2^32 elements)
__device__ __forceinline__ FIdx fdecode(const KGrid &g, long idx) {
  FIdx f;
  const unsigned u = (unsigned)idx, nr = (unsigned)g.nrchopdim, nt = (unsigned)g.ntchopdim;
  const unsigned r = u / nr;
  f.n = (int)(u - r * nr);
  f.k = (int)(r / nt);
  f.mm = (int)(r - (unsigned)f.k * nt);
  return f;
}
__device__ __forceinline__ size_t fcol(const KGrid &g, int mm, int k) {
  return (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
}

// tri-diagonal band row n of mode mm applied to column x (truncated at nn)
// (r: the table's row n, in memory or already in registers)
__device__ __forceinline__ double2 band3_r(const double *r, const double2 *x, int n, int nn) {
  double2 o = axc(r[1], x[n]);
  if (n >= 1) o = add2(o, axc(r[0], x[n - 1]));
  if (n + 1 < nn) o = add2(o, axc(r[2], x[n + 1]));
  return o;
}
__device__ __forceinline__ double2 band3_row(const double *bt, const double2 *x, int n, int nn) {
  return band3_r(bt + 3 * n, x, n, nn);
}
// penta-diagonal row
__device__ __forceinline__ double2 band5_r(const double *r, const double2 *x, int n, int nn) {
  double2 o = axc(r[2], x[n]);
  if (n >= 2) o = add2(o, axc(r[0], x[n - 2]));
  if (n >= 1) o = add2(o, axc(r[1], x[n - 1]));
  if (n + 1 < nn) o = add2(o, axc(r[3], x[n + 1]));
  if (n + 2 < nn) o = add2(o, axc(r[4], x[n + 2]));
  return o;
}
__device__ __forceinline__ double2 band5_row(const double *bt, const double2 *x, int n, int nn) {
  return band5_r(bt + 5 * n, x, n, nn);
}
// A column kernel's thread handles rows n = threadIdx.x + 256 j.  The table
// rows of its first two (j = 0, 1) are loaded into registers together with
// the column itself, before the block's first barrier, so that a block waits
// out one memory round trip instead of one per stage; rows from 512 on
// (NRCHOPDIM > 512) read the tables in place.
template <int NT>
struct RowTables {
  double r[2][NT];
  __device__ __forceinline__ void load(const double *t, int nrows) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int n = threadIdx.x + 256 * j;
#pragma unroll
      for (int e = 0; e < NT; e++) r[j][e] = n < nrows ? t[(size_t)n * NT + e] : 0.0;
    }
  }
};
// f(n, row n of table 1, row n of table 2) for the thread's rows n < nrows
// (t1, t2: the tables T1, T2 hold the first two rows of; NT is recovered
// from the size of a held row)
#define FOR_MY_ROWS(nrows, T1, t1, T2, t2, f)                                       \
  do {                                                                              \
    const int n0_ = threadIdx.x;                                                    \
    if (n0_ < (nrows)) f(n0_, T1.r[0], T2.r[0]);                                    \
    if (n0_ + 256 < (nrows)) f(n0_ + 256, T1.r[1], T2.r[1]);                         \
    for (int n_ = n0_ + 512; n_ < (nrows); n_ += 256)                               \
      f(n_, (t1) + (size_t)n_ * (sizeof(T1.r[0]) / 8), (t2) + (size_t)n_ * (sizeof(T2.r[0]) / 8)); \
  } while (0)

__host__ __device__ static inline long fff_count(const KGrid &g) { return (long)g.nrchopdim * g.ntchopdim * g.nxchopdim; }
static int nblocks(long total) {
  long b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  return (int)(b > 0 ? b : 1);
}
// one block per (m, k) column of an FFF field
static inline dim3 column_grid(const KGrid &g) { return dim3(g.ntchopdim, g.nxchopdim); }
#define GRID_STRIDE(total) for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (total); idx += (long)gridDim.x * blockDim.x)

// ---------------------------------------------------------------- basics
__global__ void k_zero(double2 *a, long n) {
  GRID_STRIDE(n) a[idx] = zero2();
}
__global__ void k_copy(double2 *d, const double2 *s, long n) {
  GRID_STRIDE(n) d[idx] = s[idx];
}
__global__ void k_negate(double2 *a, long n) {
  GRID_STRIDE(n) a[idx] = make_double2(-a[idx].x, -a[idx].y);
}
__global__ void k_scale(double2 *a, double s, long n) {
  GRID_STRIDE(n) a[idx] = axc(s, a[idx]);
}
__global__ void k_set_ln(double *dst, const double *src, double scale, double add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = (src ? src[0] * scale : 0.0) + add;
}

int launch_zero(mlg_ctx *c, double2 *a, long n) {
  hipLaunchKernelGGL(k_zero, dim3(nblocks(n)), dim3(256), 0, c->stream, a, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_copy(mlg_ctx *c, double2 *dst, const double2 *src, long n) {
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(256), 0, c->stream, dst, src, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
// COPY0 of an FFF field in one pass: E inside the truncation, 0 outside
// (the chop of the copy), and LN; one 64-thread block per (m, k) column
__global__ void __launch_bounds__(64) k_copy0_fff(KGrid g, double2 *dst, double *dst_ln, const double2 *src,
                                                  const double *src_ln) {
  const int mm = blockIdx.x, k = blockIdx.y;
  const int nn = mm < g.ntchop ? s_nrchops(g.nrchop, mm * g.minc) : 0;
  const size_t col = (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
  for (int n = threadIdx.x; n < g.nrchopdim; n += 64) dst[col + n] = n < nn ? src[col + n] : zero2();
  if (mm == 0 && k == 0 && threadIdx.x == 0) dst_ln[0] = src_ln[0] * 1.0 + 0.0;
}
int launch_copy0_fff(mlg_ctx *c, double2 *dst, double *dst_ln, const double2 *src, const double *src_ln) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_copy0_fff, dim3(g.ntchopdim, g.nxchopdim), dim3(64), 0, c->stream, g, dst, dst_ln, src, src_ln);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_negate(mlg_ctx *c, double2 *a, long n) {
  hipLaunchKernelGGL(k_negate, dim3(nblocks(n)), dim3(256), 0, c->stream, a, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_scale(mlg_ctx *c, double2 *a, double s, long n) {
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(256), 0, c->stream, a, s, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
struct LnList {
  double *p[32];
  int n;
};
__global__ void k_zero_lns(LnList l) {
  if ((int)threadIdx.x < l.n) l.p[threadIdx.x][0] = 0.0;
}
int launch_zero_lns(mlg_ctx *c, double *const *ptrs, int n) {
  for (int i0 = 0; i0 < n; i0 += 32) {
    LnList l;
    l.n = n - i0 < 32 ? n - i0 : 32;
    for (int i = 0; i < l.n; i++) l.p[i] = ptrs[i0 + i];
    hipLaunchKernelGGL(k_zero_lns, dim3(1), dim3(32), 0, c->stream, l);
    MLG_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

int launch_set_ln(mlg_ctx *c, double *dst, const double *src, double scale, double add) {
  hipLaunchKernelGGL(k_set_ln, dim3(1), dim3(64), 0, c->stream, dst, src, scale, add);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- CHOPDO
// src/mod_scalar3.f90:213-257.  PFP keeps only m < NTCHOPDIM meaningful.
// one 64-thread block per (m, k) column: zero rows n >= NRCHOPS(m), and the
// whole column for m >= NTCHOP (only the zeroed positions are visited)
// The value-at-infinity fix-up of one m = 0 column (one 64-thread block):
// A(0) -= sum_n AT1(n) A(n) / AT1(0), the products formed by all lanes, the
// sum by lane 0 in ascending n
__device__ void fix_inf_column(const KGrid &g, const double *at1, double2 *col, double2 *p) {
  for (int n = threadIdx.x; n < g.nrchop; n += 64) p[n] = axc(at1[n], col[n]);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 s = zero2();
#pragma unroll 8
    for (int n = 0; n < g.nrchop; n++) s = add2(s, p[n]);
    const double2 v0 = col[0];
    col[0] = sub2(v0, make_double2(s.x / g.at1_0, s.y / g.at1_0));
  }
}

// CHOPDO of FFF fields (grid layer z = field z of the list); with fix set,
// also the value-at-infinity fix-up of each field (rows below NRCHOP of the
// m = 0 columns, which the chop does not touch)
__global__ void __launch_bounds__(64) k_chopdo_fff(KGrid g, FieldList fl, int fix, const double *at1) {
  extern __shared__ __align__(16) double2 fsp[];
  double2 *a = fl.e[blockIdx.z];
  const int mm = blockIdx.x, k = blockIdx.y;
  const int nn = mm < g.ntchop ? s_nrchops(g.nrchop, mm * g.minc) : 0;
  double2 *col = a + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
  for (int n = (nn > 0 ? nn : 0) + threadIdx.x; n < g.nrchopdim; n += 64) col[n] = zero2();
  if (fix && mm == 0 && g.nrchop >= 1) fix_inf_column(g, at1, col, fsp);
}
__global__ void k_chopdo_rows(KGrid g, double2 *a, int n2, int n3, int stride2, int zero_slot) {
  // rows NR..NDIMR-1 for every (j < n2, z < n3), then (optionally) every
  // row of theta slot `zero_slot`; only the zeroed positions are visited
  const unsigned pad = (unsigned)(g.ndimr - g.nr);
  const unsigned ta = pad * (unsigned)n2 * (unsigned)n3;
  const unsigned tb = zero_slot >= 0 ? (unsigned)g.ndimr * (unsigned)n3 : 0u;
  GRID_STRIDE((long)ta + tb) {
    const unsigned u = (unsigned)idx;
    size_t id;
    if (u < ta) {
      const unsigned q = u / pad, r = g.nr + (u - q * pad);
      const unsigned z = q / (unsigned)n2, j = q - z * (unsigned)n2;
      id = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)stride2 * z);
    } else {
      const unsigned v = u - ta, z = v / (unsigned)g.ndimr, r = v - z * (unsigned)g.ndimr;
      id = (size_t)r + (size_t)g.ndimr * ((size_t)zero_slot + (size_t)stride2 * z);
    }
    a[id] = zero2();
  }
}

int launch_chopdo(mlg_ctx *c, double2 *a, int space) {
  KGrid g = kgrid(c);
  auto rows = [&](int n2, int n3, int stride2, int zero_slot) {
    long tot = (long)(g.ndimr - g.nr) * n2 * n3 + (zero_slot >= 0 ? (long)g.ndimr * n3 : 0);
    if (tot > 0)
      hipLaunchKernelGGL(k_chopdo_rows, dim3(nblocks(tot)), dim3(256), 0, c->stream, g, a, n2, n3, stride2, zero_slot);
  };
  if (space == MLG_FFF) {
    FieldList fl;
    fl.e[0] = a;
    hipLaunchKernelGGL(k_chopdo_fff, dim3(g.ntchopdim, g.nxchopdim, 1), dim3(64), 0, c->stream, g, fl, 0, nullptr);
  } else if (space == MLG_PFF) {
    rows(g.ntchopdim, g.nxchopdim, g.ntchopdim, -1);
  } else if (space == MLG_PPP) {
    rows(g.ndimth, g.ndimx, g.ndimth, g.ndimth > g.nth ? g.nth : -1);
  } else {  // PFP: rows >= NR of the live slots m < NTCHOPDIM
    rows(g.ntchopdim, g.ndimx, g.ndimth, -1);
  }
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_chopdo_fff_n(mlg_ctx *c, int n, double2 *const *a, bool fix) {
  if (n < 1 || n > MAX_FIELD_LIST) return MLG_E_ARG;
  KGrid g = kgrid(c);
  FieldList fl;
  for (int i = 0; i < n; i++) fl.e[i] = a[i];
  const size_t lds = fix ? sizeof(double2) * (size_t)(g.nrchop > 0 ? g.nrchop : 1) : 0;
  hipLaunchKernelGGL(k_chopdo_fff, dim3(g.ntchopdim, g.nxchopdim, n), dim3(64), lds, c->stream, g, fl, fix ? 1 : 0,
                     c->d.at1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- DELSQH / IDELSQH
// src/mod_legops.f90:76-239
__global__ void k_delsqh(KGrid g, const double2 *a, const double *ln_a, double2 *b, double *ln_b) {
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    int mv = f.mm * g.minc, nn = s_nrchops(g.nrchop, mv);
    double2 v = zero2();
    if (f.n < nn && f.mm < g.ntchop) {
      int N = mv + f.n;
      double cfac = (double)(N * (N + 1)) / g.ell2;
      double2 x = a[idx];
      v = make_double2(-x.x * cfac, -x.y * cfac);
    }
    if (idx == 0) v = make_double2(ln_a[0] / g.ell2 / g.norm0, 0.0);
    b[idx] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ln_b[0] = 0.0;
}

__global__ void k_idelsqh(KGrid g, const double2 *b, double2 *a, double *ln_a) {
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    int mv = f.mm * g.minc, nn = s_nrchops(g.nrchop, mv);
    double2 v = zero2();
    if (f.n < nn && f.mm < g.ntchop) {
      int N = mv + f.n;
      if (N != 0) {
        double cfac = g.ell2 / (double)(N * (N + 1));
        double2 x = b[idx];
        v = make_double2(-x.x * cfac, -x.y * cfac);
      }
    }
    a[idx] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ln_a[0] = b[0].x * g.norm0 * g.ell2;
}

int launch_delsqh(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *b, double *ln_b) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_delsqh, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, a, ln_a, b, ln_b);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_idelsqh(mlg_ctx *c, const double2 *b, double2 *a, double *ln_a) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_idelsqh, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, b, a, ln_a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- MULXM / XXDX
// which: 0 = MULXM (:371-428), 1 = XXDX (:680-748, adds LN/N0, LN/N1 at m0k0)
__global__ void k_band3(KGrid g, const double *bxm, const double *bxxdx, int which, const double2 *a, const double *ln_a,
                        double2 *b, double *ln_b) {
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    int mv = f.mm * g.minc, nn = s_nrchops(g.nrchop, mv);
    double2 v = zero2();
    if (f.n < nn && f.mm < g.ntchop) {
      const double *bt = (which ? bxxdx : bxm) + (size_t)f.mm * g.nband * 3;
      v = band3_row(bt, a + fcol(g, f.mm, f.k), f.n, nn);
    }
    if (which == 1 && f.mm == 0 && f.k == 0) {
      if (f.n == 0) v.x = v.x + ln_a[0] / g.norm0;
      if (f.n == 1) v.x = v.x + ln_a[0] / g.norm1;
    }
    b[idx] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ln_b[0] = 0.0;
}

int launch_band3(mlg_ctx *c, int which, const double2 *a, const double *ln_a, double2 *b, double *ln_b) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_band3, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, c->d.bxm, c->d.bxxdx, which, a, ln_a, b,
                     ln_b);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- DEL2
// (:751-820): penta band - k^2, plus the log image 4/3, -2, 2/3 at m0k0.
__device__ __forceinline__ double2 del2_row(const KGrid &g, const double *bd, const double *ak, const double2 *x, double ln,
                                            int mm, int k, int n, int nn) {
  double2 o = band5_row(bd + (size_t)mm * g.nband * 5, x, n, nn);
  double a = ak[mm + (size_t)g.ntchopdim * k];
  o = sub2(o, axc(a * a, x[n]));
  if (mm == 0 && k == 0) {
    if (n == 0) o.x = o.x + 4.0 / 3 * ln / g.ell2 / g.norm0;
    if (n == 1) o.x = o.x - 2.0 * ln / g.ell2 / g.norm1;
    if (n == 2) o.x = o.x + 2.0 / 3 * ln / g.ell2 / g.norm2;
  }
  return o;
}

__global__ void k_del2(KGrid g, const double *bd, const double *ak, const double2 *a, const double *ln_a, double2 *b,
                       double *ln_b) {
  const double ln = ln_a[0];
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    int mv = f.mm * g.minc, nn = s_nrchops(g.nrchop, mv);
    double2 v = zero2();
    if (f.n < nn && f.mm < g.ntchop) v = del2_row(g, bd, ak, a + fcol(g, f.mm, f.k), ln, f.mm, f.k, f.n, nn);
    b[idx] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ln_b[0] = 0.0;
}

int launch_del2(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *b, double *ln_b) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_del2, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, c->d.bdel2h, c->d.ak, a, ln_a, b, ln_b);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// WK%LN = -ALP*A%LN, and DEL2's top-mode check of AO at NRCHOP+2
// (one thread of each job)
__device__ __forceinline__ void visc2_ln_check(const KGrid &g, int nrchop2, const Visc2Job &J, double alp, int *err) {
  J.ln_wk[0] = J.ln_a[0] * -alp + 0.0;
  if (g.ntchopdim >= 2) {
    const int nt = s_nrchops(nrchop2, g.minc);
    if (nt >= 1 && g.ntchop > 1 && nt - 1 < s_nrchops(g.nrchop, g.minc)) {
      double2 tv = J.ao[(size_t)(nt - 1) + (size_t)g.nrchopdim * 1];
      if (hypot(tv.x, tv.y) > 1.0e-15) atomicCAS(err, 0, 3);
    }
  }
}

// VISC2's right-hand side in one pass (mod_march.f90:1579-1636):
//   CHOPSET(2); WK = DEL2(AO); CHOPSET(-2); CHOPDO(WK); WK = -WK - ALP*A;
//   WK%LN = -ALP*A%LN
// with DEL2 evaluated on the truncation NRCHOP+2 (nrchop2) at the positions
// the base truncation keeps, and DEL2's top-mode check of AO at NRCHOP+2.
// k_visc2_rhs_il writes WK row-interleaved over the systems: row n of
// system s = m + NTCHOPDIM k at wk[n * NSYS + s] (NSYS = NTCHOPDIM *
// NXCHOPDIM), the layout in which the band substitution, one thread per
// system, reads it coalesced.  A block takes VR rows of 64 consecutive
// systems (grid: systems, row chunks, jobs): four threads per column form
// VR/4 consecutive rows each (DEL2's row as del2_row forms it, then
// -WK - ALP*A), and the tile goes through LDS, each wave writing whole rows
// of the 64 systems.
template <int VR>
__global__ void __launch_bounds__(256) k_visc2_rhs_il(KGrid g, int nrchop2, const double *__restrict__ bd,
                                                      const double *__restrict__ ak, Visc2Jobs jobs, int *err) {
  constexpr int RPT = VR / 4;
  static_assert(RPT * 5 % 2 == 0, "row coefficients in pairs");
  __shared__ double2 so[64][VR + 1];
  const Visc2Job &J = jobs.j[blockIdx.z];
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long s0 = (long)blockIdx.x * 64;
  const int t = threadIdx.x;
  const double ln = J.ln_ao[0], alp = J.alp;
  // column role: column cs of the block, rows sub*RPT .. sub*RPT+RPT-1 of the chunk
  const int cs = t >> 2, sub = t & 3;
  const long ss = s0 + cs;
  const bool svalid = ss < nsys;
  const int ms = svalid ? (int)(ss % g.ntchopdim) : 0, ks = svalid ? (int)(ss / g.ntchopdim) : 0;
  const int nn = (svalid && ms < g.ntchop) ? s_nrchops(g.nrchop, ms * g.minc) : 0;  // AO rows >= NRCHOPS read as zero
  const int nn2 = s_nrchops(nrchop2, ms * g.minc);
  const double kv = ak[ms + (size_t)g.ntchopdim * ks], k2 = kv * kv;
  const bool m0k0 = ms == 0 && ks == 0;
  const size_t cb = svalid ? (size_t)g.nrchopdim * (size_t)ss : 0;  // column (m, k) = system ss
  const double *bdm = bd + (size_t)ms * g.nband * 5;
  // write role: system c = t & 63, rows rq*RPT .. of the chunk
  const int c = t & 63, rq = t >> 6;
  const long sys = s0 + c;
  {
    const int r0 = blockIdx.y * VR;
    const int nb = r0 + sub * RPT;
    // every load of the tile first, at clamped (in-bounds) rows, so that they
    // are all in flight together; the values outside the column's rows are
    // replaced by zeros afterwards
    const int nlast = g.nrchopdim - 1;
    // (a tile wholly at or below the column's truncation reads no AO rows
    // and no band rows)
    double2 x[RPT + 4], an[RPT];
    double rb[RPT][5];
#pragma unroll
    for (int i = 0; i < RPT; i++) an[i] = J.a[cb + min(nb + i, nlast)];
    if (nb - 2 < nn) {
#pragma unroll
      for (int i = 0; i < RPT + 4; i++) x[i] = J.ao[cb + min(max(nb - 2 + i, 0), nlast)];
      const double *b0 = bdm + (size_t)nb * 5;  // the RPT rows' 5 RPT coefficients are contiguous
      if (nb + RPT <= g.nband && (reinterpret_cast<uintptr_t>(b0) & 15) == 0) {
        const double2 *bv = reinterpret_cast<const double2 *>(b0);
#pragma unroll
        for (int q = 0; q < RPT * 5 / 2; q++) {
          const double2 v = bv[q];
          rb[(2 * q) / 5][(2 * q) % 5] = v.x;
          rb[(2 * q + 1) / 5][(2 * q + 1) % 5] = v.y;
        }
      } else {
#pragma unroll
        for (int i = 0; i < RPT; i++) {
          const int n = min(nb + i, nlast);  // DEL2's table holds NRCHOPDIM+2 rows per mode
#pragma unroll
          for (int e = 0; e < 5; e++) rb[i][e] = bdm[(size_t)n * 5 + e];
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < RPT + 4; i++) x[i] = zero2();
#pragma unroll
      for (int i = 0; i < RPT; i++)
#pragma unroll
        for (int e = 0; e < 5; e++) rb[i][e] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < RPT + 4; i++) {
      const int n = nb - 2 + i;
      if (!(n >= 0 && n < nn)) x[i] = zero2();
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int n = nb + i;
      double2 v = zero2();
      if (n < nn) {  // band5_row, then -k^2 and the m0k0 log image (del2_row)
        const double *r = rb[i];
        double2 o = axc(r[2], x[i + 2]);
        if (n >= 2) o = add2(o, axc(r[0], x[i]));
        if (n >= 1) o = add2(o, axc(r[1], x[i + 1]));
        if (n + 1 < nn2) o = add2(o, axc(r[3], x[i + 3]));
        if (n + 2 < nn2) o = add2(o, axc(r[4], x[i + 4]));
        o = sub2(o, axc(k2, x[i + 2]));
        if (m0k0) {
          if (n == 0) o.x = o.x + 4.0 / 3 * ln / g.ell2 / g.norm0;
          if (n == 1) o.x = o.x - 2.0 * ln / g.ell2 / g.norm1;
          if (n == 2) o.x = o.x + 2.0 / 3 * ln / g.ell2 / g.norm2;
        }
        v = o;
      }
      const double2 a = (svalid && n < g.nrchopdim) ? an[i] : zero2();
      so[cs][sub * RPT + i] = sub2(make_double2(-v.x, -v.y), axc(alp, a));
    }
    __syncthreads();
    if (sys < nsys) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int rl = rq * RPT + q, n = r0 + rl;
        if (n < g.nrchopdim) J.wk[(size_t)n * nsys + sys] = so[c][rl];
      }
    }
  }
  if (s0 == 0 && blockIdx.y == 0 && t == 0) visc2_ln_check(g, nrchop2, J, alp, err);
}

// The same right-hand sides written in the FFF columns, the layout that the
// partitioned substitution stages with coalesced loads.  A block takes VR
// rows of up to 64 columns of one m (k = k0 .. k0+63), so DEL2's band rows
// of that m are loaded once for the block; the AO rows r0-2 .. r0+VR+1 of
// its columns go through LDS, loaded with consecutive lanes on consecutive
// rows of a column.  The grid is 1D and XCD-ordered with the row chunks of
// one column group consecutive, so a chunk's halo rows come from the L2
// that its neighbours filled.  Four threads per column form VR/4
// consecutive rows each by k_visc2_rhs_il's operations; each wave writes VR
// consecutive rows of 64 / VR columns, adding -ALP*A on the way.
template <int VR>
__global__ void __launch_bounds__(256) k_visc2_rhs_col(KGrid g, int nrchop2, const double *__restrict__ bd,
                                                       const double *__restrict__ ak, Visc2Jobs jobs, int nrc, int *err) {
  constexpr int RPT = VR / 4, XR = VR + 4, XP = XR + 1;
  static_assert(64 * XR % 256 == 0 && 64 * VR % 256 == 0, "whole tile rounds per block");
  __shared__ double2 sx[64 * XP];
  __shared__ double sb[VR * 5];
  const int nkc = (g.nxchopdim + 63) / 64;
  const int lid = xcd_logical(nrc * g.ntchopdim * nkc);
  if (lid < 0) return;
  const Visc2Job &J = jobs.j[blockIdx.y];
  const int rc = lid % nrc, cg = lid / nrc;
  const int ms = cg % g.ntchopdim, k0 = (cg / g.ntchopdim) * 64;
  const int r0 = rc * VR, t = threadIdx.x, nk = min(64, g.nxchopdim - k0), nlast = g.nrchopdim - 1;
  const int nn = ms < g.ntchop ? s_nrchops(g.nrchop, ms * g.minc) : 0;  // AO rows >= NRCHOPS read as zero
  const int nn2 = s_nrchops(nrchop2, ms * g.minc);
  const size_t kstride = (size_t)g.nrchopdim * g.ntchopdim;             // column (m, k) to (m, k + 1)
  const size_t cb0 = (size_t)g.nrchopdim * ms + kstride * (size_t)k0;  // column (m, k0)
  const double ln = J.ln_ao[0], alp = J.alp;
  const bool rows = r0 - 2 < nn;  // the chunk holds rows of the truncation (uniform over the block)
  if (rows) {
    const double *bdm = bd + (size_t)ms * g.nband * 5;  // DEL2's table holds NRCHOPDIM+2 rows per mode
    if (t < VR * 5) sb[t] = bdm[(size_t)min(r0 + t / 5, nlast) * 5 + t % 5];
#pragma unroll
    for (int q = 0; q < 64 * XR / 256; q++) {
      const int e = q * 256 + t, col = e / XR, row = e % XR, n = r0 - 2 + row;
      double2 v = zero2();
      if (col < nk && n >= 0 && n < nn) v = J.ao[cb0 + kstride * col + n];
      sx[col * XP + row] = v;
    }
  }
  __syncthreads();
  const int cs = t >> 2, sub = t & 3, ks = k0 + cs;
  double2 res[RPT];
  {
    const double kv = ak[ms + (size_t)g.ntchopdim * (cs < nk ? ks : 0)], k2 = kv * kv;
    const bool m0k0 = ms == 0 && ks == 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int rl = sub * RPT + i, n = r0 + rl;
      double2 v = zero2();
      if (rows && n < nn) {  // band5_row, then -k^2 and the m0k0 log image (del2_row)
        const double *r = sb + rl * 5;
        const double2 *x = sx + cs * XP + rl;  // x[j]: row n - 2 + j
        double2 o = axc(r[2], x[2]);
        if (n >= 2) o = add2(o, axc(r[0], x[0]));
        if (n >= 1) o = add2(o, axc(r[1], x[1]));
        if (n + 1 < nn2) o = add2(o, axc(r[3], x[3]));
        if (n + 2 < nn2) o = add2(o, axc(r[4], x[4]));
        o = sub2(o, axc(k2, x[2]));
        if (m0k0) {
          if (n == 0) o.x = o.x + 4.0 / 3 * ln / g.ell2 / g.norm0;
          if (n == 1) o.x = o.x - 2.0 * ln / g.ell2 / g.norm1;
          if (n == 2) o.x = o.x + 2.0 / 3 * ln / g.ell2 / g.norm2;
        }
        v = o;
      }
      res[i] = v;
    }
  }
  __syncthreads();  // every read of the AO tile is done
#pragma unroll
  for (int i = 0; i < RPT; i++) sx[cs * XP + sub * RPT + i] = res[i];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 64 * VR / 256; q++) {
    const int e = q * 256 + t, cc = e / VR, rl = e % VR, n = r0 + rl;
    if (cc < nk && n < g.nrchopdim) {
      const size_t o = cb0 + kstride * cc + n;
      const double2 v = sx[cc * XP + rl];
      J.wk[o] = sub2(make_double2(-v.x, -v.y), axc(alp, J.a[o]));
    }
  }
  if (lid == 0 && t == 0) visc2_ln_check(g, nrchop2, J, alp, err);
}
int launch_visc2_rhs_il(mlg_ctx *c, int n, const Visc2Jobs &jobs, int il) {
  if (n < 1 || n > 3) return MLG_E_ARG;
  KGrid g = kgrid(c);
  const int nrchop2 = g.nrchop + 2;
  if (nrchop2 > g.nrchopdim) return 1;  // CHOPSET: NRCHOP TOO LARGE (mod_scalar3.f90:183-190)
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  constexpr int VR = 16;
  const int nrc = (g.nrchopdim + VR - 1) / VR;
  if (il) {
    const dim3 grid((unsigned)((nsys + 63) / 64), (unsigned)nrc, n);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_visc2_rhs_il<VR>), grid, dim3(256), 0, c->stream, g, nrchop2, c->d.bdel2h,
                       c->d.ak, jobs, c->d.err);
  } else {
    const long nblk = (long)nrc * g.ntchopdim * ((g.nxchopdim + 63) / 64);
    if (nblk > INT_MAX / 2) return MLG_E_ARG;
    const dim3 grid((unsigned)xcd_grid((int)nblk), n);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_visc2_rhs_col<VR>), grid, dim3(256), 0, c->stream, g, nrchop2, c->d.bdel2h,
                       c->d.ak, jobs, nrc, c->d.err);
  }
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// DEL2's exactness check (:778-790): the top mode of m=1 at k=0 must vanish.
__global__ void k_check_top_m1(KGrid g, const double2 *a, int *err) {
  if (g.ntchopdim < 2) return;
  int nn = s_nrchops(g.nrchop, g.minc);
  if (nn < 1) return;
  double2 v = a[(size_t)(nn - 1) + (size_t)g.nrchopdim * 1];
  if (hypot(v.x, v.y) > 1.0e-15) atomicCAS(err, 0, 3);
}
int launch_check_top_m1(mlg_ctx *c, const double2 *a) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_check_top_m1, dim3(1), dim3(1), 0, c->stream, g, a, c->d.err);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- GRAD
// (:1727-1786): (r d/dr b, d/dtheta b, d/dz b) = (XXDX b, i m b, i k b)
__global__ void __launch_bounds__(256) k_grad(KGrid g, const double *bxxdx, const double *ak, const double2 *a,
                                             const double *ln_a, double2 *rr, double *ln_rr, double2 *rp, double *ln_rp,
                                             double2 *z, double *ln_z, double *zl0, double *zl1) {
  extern __shared__ __align__(16) double2 csm[];
  const double ln = ln_a[0];
  const int mm = blockIdx.x, k = blockIdx.y;
  const int mv = mm * g.minc, nn = mm < g.ntchop ? s_nrchops(g.nrchop, mv) : 0;
  const size_t cb = fcol(g, mm, k);
  const double *xx = bxxdx + (size_t)mm * g.nband * 3;
  RowTables<3> TX;
  TX.load(xx, nn);
  const double kv = ak[mm + (size_t)g.ntchopdim * k];
  for (int n = threadIdx.x; n < nn; n += blockDim.x) csm[n] = a[cb + n];
  __syncthreads();
  auto out = [&](int n, const double *rx, const double *) {
    double2 vr = zero2(), vp = zero2(), vz = zero2();
    if (n < nn) {
      vr = band3_r(rx, csm, n, nn);
      double2 x = csm[n];
      vp = ix((double)mv, x);
      vz = ix(kv, x);
    }
    if (mm == 0 && k == 0) {
      if (n == 0) vr.x = vr.x + ln / g.norm0;
      if (n == 1) vr.x = vr.x + ln / g.norm1;
    }
    rr[cb + n] = vr;
    if (rp) {  // null rp, z: the r d/dr component only
      rp[cb + n] = vp;
      z[cb + n] = vz;
    }
  };
  FOR_MY_ROWS(g.nrchopdim, TX, xx, TX, xx, out);
  if (mm == 0 && k == 0 && threadIdx.x == 0) {
    ln_rr[0] = 0.0;
    if (rp) {
      ln_rp[0] = 0.0;
      ln_z[0] = 0.0;
    }
    if (zl0) zl0[0] = 0.0;
    if (zl1) zl1[0] = 0.0;
  }
}
int launch_grad(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *rr, double *ln_rr, double2 *rp, double *ln_rp,
                double2 *z, double *ln_z, double *zl0, double *zl1) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_grad, column_grid(g), dim3(256), sizeof(double2) * (size_t)g.nrchopdim, c->stream, g, c->d.bxxdx,
                     c->d.ak, a, ln_a, rr, ln_rr, rp, ln_rp, z, ln_z, zl0, zl1);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// Column kernels: one block per (m, k) column.  The column's inputs and the
// intermediate rows (DELSQH, MULXM, DEL2) are staged in LDS, and every row
// is formed by the operations of the reference's row recurrences, reading
// LDS instead of memory.
// u_z-like rows -(1-x)^2 [(1-x)^-2 del_h^2 s] as two truncated MULXM passes
// over DELSQH(s) (with s's log image in row 0 of m0k0).
// W[l] = MULXM row l of DELSQH(s), the DELSQH rows formed where they are read
// row l of W from the MULXM table row r (= bxm row l of mode mm)
__device__ __forceinline__ double2 uz_stage_r(const KGrid &g, const double *r, const double2 *s, double ln, int mm, int k,
                                              int l, int nn) {
  const int mv = mm * g.minc;
  auto dsq = [&](int j) -> double2 {
    const int N = mv + j;
    const double2 x = s[j];
    const double cf = (double)(N * (N + 1)) / g.ell2;
    double2 v = make_double2(-x.x * cf, -x.y * cf);
    if (mm == 0 && k == 0 && j == 0) v = make_double2(ln / g.ell2 / g.norm0, 0.0);
    return v;
  };
  double2 o = axc(r[1], dsq(l));
  if (l >= 1) o = add2(o, axc(r[0], dsq(l - 1)));
  if (l + 1 < nn) o = add2(o, axc(r[2], dsq(l + 1)));
  return o;
}
__device__ void uz_stage_direct(const KGrid &g, const double *bxm, const double2 *s, double ln, int mm, int k, int nn,
                                double2 *W) {
  const double *xt = bxm + (size_t)mm * g.nband * 3;
  for (int l = threadIdx.x; l < nn; l += blockDim.x) W[l] = uz_stage_r(g, xt + 3 * l, s, ln, mm, k, l, nn);
}
// row n of the u_z-like result: -(MULXM row n of W)
__device__ __forceinline__ double2 uz_from_w_r(const double *r, const double2 *W, int n, int nn) {
  const double2 o = band3_r(r, W, n, nn);
  return make_double2(-o.x, -o.y);
}
__device__ __forceinline__ double2 uz_from_w(const KGrid &g, const double *bxm, const double2 *W, int mm, int n, int nn) {
  return uz_from_w_r(bxm + (size_t)mm * g.nband * 3 + 3 * n, W, n, nn);
}
// ---------------------------------------------------------------- PC2VEL
// (:1789-1853): r u_r = i m psi + i k XXDX chi ; r u_th = -XXDX psi - m k chi ;
// u_z = -(1-x) (1-x) DELSQH chi.
// HIST: the velocities only update the pressure-gradient histories H (in
// rur, rup, uz): H = H - u (hmode 4) or H = a (H - u) (hmode 3), the
// updates of the ETD steps, and are not stored (k_pc2vel_col below)
template <bool HIST>
__global__ void k_pc2vel_col(KGrid g, const double *bxm, const double *bxxdx, const double *ak, const double2 *psi,
                             const double *ln_psi, const double2 *chi, const double *ln_chi, double2 *rur, double2 *rup,
                             double2 *uz, double *l1, double *l2, double *l3, int hmode, double ha, VelTriple base,
                             const double *fat1, int fnr);
int launch_pc2vel(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi, double2 *rur,
                  double2 *rup, double2 *uz, double *ln_rur, double *ln_rup, double *ln_uz, int fix_nrchop) {
  KGrid g = kgrid(c);
  const bool fix = fix_nrchop >= 1;
  if (fix && fix_nrchop > g.nrchop) return MLG_E_ARG;  // the fix-up reads staged rows only
  hipLaunchKernelGGL(k_pc2vel_col<false>, column_grid(g), dim3(256), 3 * sizeof(double2) * (size_t)g.nrchopdim,
                     c->stream, g, c->d.bxm, c->d.bxxdx, c->d.ak, psi, ln_psi, chi, ln_chi, rur, rup, uz, ln_rur, ln_rup,
                     ln_uz, 0, 0.0, VelTriple{{nullptr, nullptr, nullptr}}, fix ? c->d.at1 : nullptr, fix_nrchop);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_pc2vel_hist(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                       double2 *hr, double2 *hp, double2 *hz, int hmode, double ha, const VelTriple &base) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_pc2vel_col<true>, column_grid(g), dim3(256), 3 * sizeof(double2) * (size_t)g.nrchopdim, c->stream, g,
                     c->d.bxm, c->d.bxxdx, c->d.ak, psi, ln_psi, chi, ln_chi, hr, hp, hz, nullptr, nullptr, nullptr, hmode, ha,
                     base, nullptr, 0);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- PC2VOR
// (:1856-1919): with Q = DEL2 chi,
//   r om_r = -i m Q + i k XXDX psi ; r om_th = XXDX Q - k m psi ;
//   om_z = -(1-x)(1-x) DELSQH psi.
__global__ void __launch_bounds__(256) k_pc2vor(KGrid g, const double *bxm, const double *bxxdx, const double *bd,
                                               const double *ak, const double2 *psi, const double *ln_psi,
                                               const double2 *chi, const double *ln_chi, double2 *ror, double2 *rop,
                                               double2 *oz, double *l1, double *l2, double *l3) {
  extern __shared__ __align__(16) double2 csm[];
  const int mm = blockIdx.x, k = blockIdx.y;
  const int mv = mm * g.minc, nn = mm < g.ntchop ? s_nrchops(g.nrchop, mv) : 0;
  const size_t cb = fcol(g, mm, k);
  double2 *P = csm, *X = P + g.nrchopdim, *Q = X + g.nrchopdim, *W = Q + g.nrchopdim;
  const double lp = ln_psi[0], lc = ln_chi[0];
  for (int n = threadIdx.x; n < nn; n += blockDim.x) {
    P[n] = psi[cb + n];
    X[n] = chi[cb + n];
  }
  __syncthreads();
  for (int n = threadIdx.x; n < nn; n += blockDim.x) Q[n] = del2_row(g, bd, ak, X, lc, mm, k, n, nn);
  uz_stage_direct(g, bxm, P, lp, mm, k, nn, W);
  __syncthreads();
  const double *xx = bxxdx + (size_t)mm * g.nband * 3;
  const double kv = ak[mm + (size_t)g.ntchopdim * k];
  for (int n = threadIdx.x; n < g.nrchopdim; n += blockDim.x) {
    double2 vr = zero2(), vp = zero2(), vz = zero2();
    if (n < nn) {
      const double2 q = Q[n];
      double2 xp = band3_row(xx, P, n, nn);
      if (mm == 0 && k == 0) {
        if (n == 0) xp.x = xp.x + lp / g.norm0;
        if (n == 1) xp.x = xp.x + lp / g.norm1;
      }
      double2 imq = ix((double)mv, q);
      vr = add2(make_double2(-imq.x, -imq.y), ix(kv, xp));
      // XXDX of Q (Q carries no log term)
      const double *r = xx + 3 * n;
      double2 xq = axc(r[1], q);
      if (n >= 1) xq = add2(xq, axc(r[0], Q[n - 1]));
      if (n + 1 < nn) xq = add2(xq, axc(r[2], Q[n + 1]));
      vp = sub2(xq, axc(kv * mv, P[n]));
      vz = uz_from_w(g, bxm, W, mm, n, nn);
    }
    ror[cb + n] = vr;
    rop[cb + n] = vp;
    oz[cb + n] = vz;
  }
  if (mm == 0 && k == 0 && threadIdx.x == 0) {
    l1[0] = 0.0;
    l2[0] = 0.0;
    l3[0] = 0.0;
  }
}
int launch_pc2vor(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi, double2 *ror,
                  double2 *rop, double2 *oz, double *ln_ror, double *ln_rop, double *ln_oz) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_pc2vor, column_grid(g), dim3(256), 4 * sizeof(double2) * (size_t)g.nrchopdim, c->stream, g, c->d.bxm,
                     c->d.bxxdx, c->d.bdel2h, c->d.ak, psi, ln_psi, chi, ln_chi, ror, rop, oz, ln_ror, ln_rop, ln_oz);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- PC2VEL + PC2VOR
// Both of one (PSI, CHI) pair in one pass, as the time steps call them
// (CHOPSET(3); PC2VEL; PC2VOR): one block per (m, k) column, the column of
// psi and chi, DEL2 chi and the two MULXM(DELSQH) stages in LDS.  Every
// output row is formed by the same operations as in k_pc2vel and k_pc2vor.
__global__ void __launch_bounds__(256) k_pc2velvor(KGrid g, const double *bxm, const double *bxxdx, const double *bd,
                                                   const double *ak, const double2 *psi, const double *ln_psi,
                                                   const double2 *chi, const double *ln_chi, VelVorOut o, int *err) {
  extern __shared__ __align__(16) double2 csm[];
  const int mm = blockIdx.x, k = blockIdx.y;
  const int mv = mm * g.minc, nn = mm < g.ntchop ? s_nrchops(g.nrchop, mv) : 0;
  const size_t cb = fcol(g, mm, k);
  double2 *P = csm, *X = P + g.nrchopdim, *Q = X + g.nrchopdim, *WP = Q + g.nrchopdim, *WX = WP + g.nrchopdim;
  const double lp = ln_psi[0], lc = ln_chi[0];
  const double *bdm = bd + (size_t)mm * g.nband * 5, *xm = bxm + (size_t)mm * g.nband * 3,
               *xx = bxxdx + (size_t)mm * g.nband * 3;
  RowTables<5> T5;
  RowTables<3> TM, TX;
  T5.load(bdm, nn);
  TM.load(xm, nn);
  TX.load(xx, nn);
  const double kv = ak[mm + (size_t)g.ntchopdim * k];
  for (int n = threadIdx.x; n < nn; n += blockDim.x) {
    P[n] = psi[cb + n];
    X[n] = chi[cb + n];
  }
  __syncthreads();
  // DEL2's exactness check of PC2VOR (k_check_top_m1): the top mode of
  // chi at m = 1, k = 0 must vanish
  if (mm == 1 && k == 0 && threadIdx.x == 0 && nn >= 1) {
    const double2 t = X[nn - 1];
    if (hypot(t.x, t.y) > 1.0e-15) atomicCAS(err, 0, 3);
  }
  // Q = DEL2 chi (del2_row), WP / WX: the MULXM(DELSQH) stages of psi / chi
  auto stage = [&](int n, const double *r5, const double *rm) {
    double2 q = band5_r(r5, X, n, nn);
    q = sub2(q, axc(kv * kv, X[n]));
    if (mm == 0 && k == 0) {
      if (n == 0) q.x = q.x + 4.0 / 3 * lc / g.ell2 / g.norm0;
      if (n == 1) q.x = q.x - 2.0 * lc / g.ell2 / g.norm1;
      if (n == 2) q.x = q.x + 2.0 / 3 * lc / g.ell2 / g.norm2;
    }
    Q[n] = q;
    WP[n] = uz_stage_r(g, rm, P, lp, mm, k, n, nn);
    WX[n] = uz_stage_r(g, rm, X, lc, mm, k, n, nn);
  };
  FOR_MY_ROWS(nn, T5, bdm, TM, xm, stage);
  __syncthreads();
  auto out = [&](int n, const double *rx, const double *rm) {
    double2 ur = zero2(), up = zero2(), uz = zero2(), wr = zero2(), wp = zero2(), wz = zero2();
    if (n < nn) {
      double2 xc = band3_r(rx, X, n, nn);
      double2 xp = band3_r(rx, P, n, nn);
      if (mm == 0 && k == 0) {
        if (n == 0) { xc.x = xc.x + lc / g.norm0; xp.x = xp.x + lp / g.norm0; }
        if (n == 1) { xc.x = xc.x + lc / g.norm1; xp.x = xp.x + lp / g.norm1; }
      }
      // velocity (PC2VEL)
      ur = add2(ix((double)mv, P[n]), ix(kv, xc));
      up = sub2(make_double2(-xp.x, -xp.y), axc(mv * kv, X[n]));
      uz = uz_from_w_r(rm, WX, n, nn);
      // vorticity (PC2VOR); XXDX of Q (Q carries no log term)
      const double2 q = Q[n];
      const double2 imq = ix((double)mv, q);
      wr = add2(make_double2(-imq.x, -imq.y), ix(kv, xp));
      const double2 xq = band3_r(rx, Q, n, nn);
      wp = sub2(xq, axc(kv * mv, P[n]));
      wz = uz_from_w_r(rm, WP, n, nn);
    }
    o.e[0][cb + n] = ur;
    o.e[1][cb + n] = up;
    o.e[2][cb + n] = uz;
    o.e[3][cb + n] = wr;
    o.e[4][cb + n] = wp;
    o.e[5][cb + n] = wz;
  };
  FOR_MY_ROWS(g.nrchopdim, TX, xx, TM, xm, out);
  if (mm == 0 && k == 0 && threadIdx.x < 6) o.ln[threadIdx.x][0] = 0.0;
}
int launch_pc2velvor(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                     const VelVorOut &o) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_pc2velvor, column_grid(g), dim3(256), 5 * sizeof(double2) * (size_t)g.nrchopdim, c->stream, g,
                     c->d.bxm, c->d.bxxdx, c->d.bdel2h, c->d.ak, psi, ln_psi, chi, ln_chi, o, c->d.err);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// PC2VEL as a column kernel (psi, chi and MULXM(DELSQH chi) of the column in
// LDS), with the HIST variant of k_pc2vel; the same operations per row.
template <bool HIST>
__global__ void __launch_bounds__(256) k_pc2vel_col(KGrid g, const double *bxm, const double *bxxdx, const double *ak,
                                                    const double2 *psi, const double *ln_psi, const double2 *chi,
                                                    const double *ln_chi, double2 *rur, double2 *rup, double2 *uz,
                                                    double *l1, double *l2, double *l3, int hmode, double ha,
                                                    VelTriple base, const double *fat1, int fnr) {
  extern __shared__ __align__(16) double2 csm[];
  const int mm = blockIdx.x, k = blockIdx.y;
  const int mv = mm * g.minc, nn = mm < g.ntchop ? s_nrchops(g.nrchop, mv) : 0;
  const size_t cb = fcol(g, mm, k);
  double2 *P = csm, *X = P + g.nrchopdim, *WX = X + g.nrchopdim;
  const double lp = ln_psi[0], lc = ln_chi[0];
  const double *xm = bxm + (size_t)mm * g.nband * 3, *xx = bxxdx + (size_t)mm * g.nband * 3;
  RowTables<3> TM, TX;
  TM.load(xm, nn);
  TX.load(xx, nn);
  const double kv = ak[mm + (size_t)g.ntchopdim * k];
  // H is read from `base` when given (the values the histories were to
  // receive just before), else from the histories themselves; the first two
  // rows' values are loaded with the column
  const double2 *hs[3] = {base.e[0] ? base.e[0] : rur, base.e[0] ? base.e[1] : rup, base.e[0] ? base.e[2] : uz};
  double2 hv[2][3];
  if (HIST) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int n = threadIdx.x + 256 * j;
#pragma unroll
      for (int q = 0; q < 3; q++) hv[j][q] = n < g.nrchopdim ? hs[q][cb + n] : zero2();
    }
  }
  // with fat1: psi and chi come straight from PROJECT and IDEL2 at the
  // truncation fnr, and this kernel is their only reader.  IDEL2's CHOPDO of
  // chi is made as the column is staged (rows >= NRCHOPS(fnr, m) read as
  // zero), and the value-at-infinity fix-ups of both m = 0 columns
  // (PROJECT :2541-2543, IDEL2 :987-992) on the staged columns:
  // A(0) -= sum_{n < fnr} AT1(n) A(n) / AT1(0), the products in parallel,
  // one thread per field summing in ascending n.
  const int nchi = !fat1 ? nn : (mm < g.ntchop ? s_nrchops(fnr, mv) : 0);
  for (int n = threadIdx.x; n < nn; n += blockDim.x) {
    P[n] = psi[cb + n];
    X[n] = n < nchi ? chi[cb + n] : zero2();
  }
  __syncthreads();
  if (fat1 && mm == 0 && fnr >= 1) {
    double2 *F = WX;  // scratch for the products until the stage below writes WX
    for (int f = 0; f < 2; f++) {
      double2 *A = f == 0 ? P : X;
      for (int n = threadIdx.x; n < fnr; n += blockDim.x) F[n] = make_double2(fat1[n] * A[n].x, fat1[n] * A[n].y);
      __syncthreads();
      if (threadIdx.x == 0) {
        double sx = 0.0, sy = 0.0;
        for (int n = 0; n < fnr; n++) {
          sx = sx + F[n].x;
          sy = sy + F[n].y;
        }
        A[0] = make_double2(A[0].x - sx / g.at1_0, A[0].y - sy / g.at1_0);
      }
      __syncthreads();
    }
  }
  auto stage = [&](int n, const double *rm, const double *) { WX[n] = uz_stage_r(g, rm, X, lc, mm, k, n, nn); };
  FOR_MY_ROWS(nn, TM, xm, TM, xm, stage);
  __syncthreads();
  auto out = [&](int n, const double *rx, const double *rm, const double2 *h) {
    double2 vr = zero2(), vp = zero2(), vz = zero2();
    if (n < nn) {
      double2 xc = band3_r(rx, X, n, nn);
      double2 xp = band3_r(rx, P, n, nn);
      if (mm == 0 && k == 0) {
        if (n == 0) { xc.x = xc.x + lc / g.norm0; xp.x = xp.x + lp / g.norm0; }
        if (n == 1) { xc.x = xc.x + lc / g.norm1; xp.x = xp.x + lp / g.norm1; }
      }
      vr = add2(ix((double)mv, P[n]), ix(kv, xc));
      vp = sub2(make_double2(-xp.x, -xp.y), axc(mv * kv, X[n]));
      vz = uz_from_w_r(rm, WX, n, nn);
    }
    const size_t i = cb + n;
    if (HIST) {
      const double2 h0 = h[0], h1 = h[1], h2 = h[2];
      if (hmode == 3) {
        rur[i] = axc(ha, sub2(h0, vr));
        rup[i] = axc(ha, sub2(h1, vp));
        uz[i] = axc(ha, sub2(h2, vz));
      } else {
        rur[i] = sub2(h0, vr);
        rup[i] = sub2(h1, vp);
        uz[i] = sub2(h2, vz);
      }
    } else {
      rur[i] = vr;
      rup[i] = vp;
      uz[i] = vz;
    }
  };
  const int n0 = threadIdx.x;
  if (n0 < g.nrchopdim) out(n0, TX.r[0], TM.r[0], hv[0]);
  if (n0 + 256 < g.nrchopdim) out(n0 + 256, TX.r[1], TM.r[1], hv[1]);
  for (int n = n0 + 512; n < g.nrchopdim; n += 256) {
    double2 hl[3];
    if (HIST) {
#pragma unroll
      for (int q = 0; q < 3; q++) hl[q] = hs[q][cb + n];
    }
    out(n, xx + (size_t)n * 3, xm + (size_t)n * 3, hl);
  }
  if (!HIST && mm == 0 && k == 0 && threadIdx.x == 0) {
    l1[0] = 0.0;
    l2[0] = 0.0;
    l3[0] = 0.0;
  }
}

// ---------------------------------------------------------------- value at infinity
// A(0,0,k) -= sum_n AT1(n) A(n,0,k) / AT1(0)   (IDELSQH :222, IDEL2 :987-992,
// IDEL2LN :1154-1159, PROJECT :2541-2543)
// One 64-thread block per k: the m = 0 column and AT1 are staged in LDS with
// coalesced loads, lane 0 sums in ascending n.
__global__ void __launch_bounds__(64) k_fix_inf(KGrid g, const double *at1, double2 *a) {
  extern __shared__ __align__(16) double2 fsp[];  // [nrchop] products
  fix_inf_column(g, at1, a + (size_t)g.nrchopdim * (size_t)g.ntchopdim * blockIdx.x, fsp);
}
int launch_fix_at_infinity(mlg_ctx *c, double2 *a) {
  KGrid g = kgrid(c);
  if (g.nrchop < 1) return 0;
  hipLaunchKernelGGL(k_fix_inf, dim3(g.nxchopdim), dim3(64), sizeof(double2) * (size_t)g.nrchop, c->stream, g, c->d.at1,
                     a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- banded solves
// LUDCMPB + LUBKSBB (src/mod_bandmat.f90:65-156,320-402) for one system per
// thread, streamed: a window of the c0+1 rows below the pivot is kept in
// registers, forward substitution follows the factorisation row by row, the
// normalised upper rows go to scratch for the back substitution.  No
// pivoting, 1/diagonal stored, same update order as the reference.
struct SolveCtx {
  KGrid g;
  const double *bd, *bxm, *ak;
  double alp;
  int op;      // 0 IDEL2, 1 IDEL2LN, 2 IHELM
  int kind;    // 0 penta, 1 IDEL2 m0k0 shifted (7-band), 2 IHELM m=0 product (7-band)
  int mm, k, np;
  double k2;
};

// original matrix row `q` as NE band entries (centre c0)
__device__ void gen_row(const SolveCtx &s, int q, double *out, int NE) {
  const KGrid &g = s.g;
  for (int b = 0; b < NE; b++) out[b] = 0.0;
  if (s.kind == 0) {
    const double *r = s.bd + ((size_t)s.mm * g.nband + q) * 5;
    out[0] = r[0];
    out[1] = r[1];
    out[2] = (r[2] - s.alp) - s.k2;
    out[3] = r[3];
    out[4] = r[4];
    if (s.mm == 0 && s.k == 0 && s.op == 1) {
      if (q == 0) out[2] = out[2] + 4.0 / 3 / g.ell2;
      if (q == 1) out[1] = out[1] - 2 * g.norm0 / g.norm1 / g.ell2;
      if (q == 2) out[0] = out[0] + 2.0 / 3 * g.norm0 / g.norm2 / g.ell2;
    }
  } else if (s.kind == 1) {
    if (q == 0) {
      out[3] = 1.0;
    } else {
      int p = q - 1;
      const double *r = s.bd + ((size_t)s.mm * g.nband + p) * 5;
      out[0] = r[0];
      out[1] = r[1];
      out[2] = r[2] - s.k2;
      out[3] = r[3];
      out[4] = r[4];
      if (p == 0) out[2] = out[2] + 4.0 / 3 / g.ell2;
      if (p == 1) out[1] = out[1] - 2 * g.norm0 / g.norm1 / g.ell2;
      if (p == 2) out[0] = out[0] + 2.0 / 3 * g.norm0 / g.norm2 / g.ell2;
    }
  } else {
    // (H5 * XM) row q, both truncated to np x np (BANBANMUL order)
    const double *h = s.bd + ((size_t)s.mm * g.nband + q) * 5;
    double h5[5] = {h[0], h[1], (h[2] - s.alp) - s.k2, h[3], h[4]};
    for (int I = 0; I < 5; I++) {
      int rr = q + I - 2;
      if (rr < 0 || rr >= s.np) continue;
      const double *xm = s.bxm + ((size_t)s.mm * g.nband + rr) * 3;
      for (int J = 0; J < 3; J++) {
        int cc = rr + J - 1;
        if (cc < 0 || cc >= s.np) continue;
        out[I + J] += h5[I] * xm[J];
      }
    }
  }
}

// Factor storage of a set: three regions over (row nn, class cls), class =
// m index + NTCHOPDIM*|k| index, NRCHOPDIM rows:
//   penta-diagonal classes (C0 = 2, KU = 2), split by substitution pass:
//     FF[(nn*ncls + cls)*3 + e]: the C0 multipliers of row nn against rows
//        nn-1-e (scaled by their pivots), then 1/pivot   (forward pass)
//     FB[(nn*ncls + cls)*2 + u]: the KU normalised upper entries (nn, nn+1+u)
//        (back substitution)
//   7-band classes (m = 0 forms): F7[(nn*ncls + cls)*FSTRIDE + e] with
//     e < C0 multipliers, C0 <= e < C0+KU upper entries, e = NE-1 1/pivot.
// Each pass reads only its part, and a wave of consecutive m reads one
// contiguous run per row.
#define FSTRIDE 7
#define FSET_DOUBLES (3 + 2 + FSTRIDE)  // per (row, class)
struct FactorView {
  double *ff, *fb, *f7;
};
__host__ __device__ __forceinline__ FactorView factor_view(const double *F, int rows, long ncls) {
  double *f = const_cast<double *>(F);
  const size_t rc = (size_t)rows * ncls;
  return FactorView{f, f + 3 * rc, f + 5 * rc};
}

template <int NE, int C0>
__device__ void band_factor(const SolveCtx &s, int nd, const FactorView &FV, long ncls, long cls) {
  constexpr int KU = NE - 1 - C0;
  double W[C0 + 1][NE];
  for (int q = 0; q <= C0; q++) {
    if (q < nd) gen_row(s, q, W[q], NE);
    else
      for (int b = 0; b < NE; b++) W[q][b] = 0.0;
  }
  for (int nn = 0; nn < nd; nn++) {
    const double piv = W[0][C0];
    for (int b = C0 + 1; b < NE; b++)
      if (nn + (b - C0) < nd) W[0][b] /= piv;
    for (int b = C0 + 1; b < NE; b++)
      for (int j = 1; j <= C0; j++) {
        int kk = (b - C0) > j ? (b - C0) : j;
        if (nn + kk < nd) W[j][b - j] -= W[0][b] * W[j][C0 - j];
      }
    for (int j = 1; j <= C0; j++)
      if (nn + j < nd) W[j][C0 - j] /= piv;
    if (NE == 5) {
      double *ff = FV.ff + ((size_t)nn * ncls + cls) * 3, *fb = FV.fb + ((size_t)nn * ncls + cls) * 2;
      for (int j = 1; j <= C0; j++) ff[j - 1] = W[0][C0 - j];
      ff[2] = 1.0 / piv;
      for (int u = 0; u < KU; u++) fb[u] = W[0][C0 + 1 + u];
    } else {
      double *fr = FV.f7 + ((size_t)nn * ncls + cls) * FSTRIDE;
      for (int j = 1; j <= C0; j++) fr[j - 1] = W[0][C0 - j];
      for (int u = 0; u < KU; u++) fr[C0 + u] = W[0][C0 + 1 + u];
      fr[NE - 1] = 1.0 / piv;
    }
    for (int q = 0; q < C0; q++)
      for (int b = 0; b < NE; b++) W[q][b] = W[q + 1][b];
    if (nn + C0 + 1 < nd) gen_row(s, nn + C0 + 1, W[C0], NE);
    else
      for (int b = 0; b < NE; b++) W[C0][b] = 0.0;
  }
}

// Forward substitution (multipliers applied to the unscaled history, then
// scaled by 1/pivot) and back substitution with the normalised U rows:
// the update order of LUBKSBB.  Rows go in register blocks of SB: all
// loads of a block are issued before its serial arithmetic, so one memory
// round trip serves SB rows of the dependency chain.
template <int NE, int C0, int SB>
__device__ void band_apply(int nd, const double2 *xin, double2 *x, const double *__restrict__ F, long ncls, long cls) {
  constexpr int KU = NE - 1 - C0;
  double2 xh[C0];
#pragma unroll
  for (int q = 0; q < C0; q++) xh[q] = zero2();
  for (int nb = 0; nb < nd; nb += SB) {
    double2 rhs[SB];
    double lm[SB][C0], rp[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nb + q;
      if (nn < nd) {
        const double *fr = F + ((size_t)nn * ncls + cls) * FSTRIDE;
        rhs[q] = xin[nn];
#pragma unroll
        for (int j = 0; j < C0; j++) lm[q][j] = fr[j];
        rp[q] = fr[NE - 1];
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nb + q;
      if (nn < nd) {
        double2 v = rhs[q];
        const int nne = nn < C0 ? nn : C0;
#pragma unroll
        for (int j = 1; j <= C0; j++)
          if (j <= nne) v = sub2(v, axc(lm[q][j - 1], xh[j - 1]));
#pragma unroll
        for (int t = C0 - 1; t >= 1; t--) xh[t] = xh[t - 1];
        xh[0] = v;
        rhs[q] = axc(rp[q], v);
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nb + q < nd) x[nb + q] = rhs[q];
  }
  double2 yh[KU];
#pragma unroll
  for (int u = 0; u < KU; u++) yh[u] = zero2();
  for (int nt = nd - 1; nt >= 0; nt -= SB) {
    double2 v0[SB];
    double um[SB][KU];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nt - q;
      if (nn >= 0) {
        const double *fr = F + ((size_t)nn * ncls + cls) * FSTRIDE;
        v0[q] = x[nn];
#pragma unroll
        for (int u = 0; u < KU; u++) um[q][u] = fr[C0 + u];
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nt - q;
      if (nn >= 0) {
        double2 v = v0[q];
        if (nn <= nd - 2) {
#pragma unroll
          for (int i = 1; i <= KU; i++)
            if (nn + i < nd) v = sub2(v, axc(um[q][i - 1], yh[i - 1]));
        }
        v0[q] = v;
#pragma unroll
        for (int u = KU - 1; u >= 1; u--) yh[u] = yh[u - 1];
        yh[0] = v;
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nt - q >= 0) x[nt - q] = v0[q];
  }
}

// band_apply for NJ right-hand sides that share one factor set: the factor
// rows are loaded once and the NJ substitutions run side by side (each with
// band_apply's arithmetic).
// The forward pass's results go to xt (row nn at xt[nn * st]: with st the
// number of systems, consecutive systems' rows are adjacent, so a wave's
// accesses coalesce), which the back substitution reads; only its results
// go to the column x.
template <int NJ, int NE, int C0, int SB>
__device__ void band_apply_n(int nd, const double2 *const (&xin)[NJ], long sin, double2 *const (&xt)[NJ],
                             double2 *const (&x)[NJ], long st, const FactorView &FV, long ncls, long cls) {
  static_assert(NE == 5 && C0 == 2, "the split factor records are penta-diagonal");
  constexpr int KU = NE - 1 - C0;
  double2 xh[NJ][C0];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
#pragma unroll
    for (int q = 0; q < C0; q++) xh[jj][q] = zero2();
  for (int nb = 0; nb < nd; nb += SB) {
    double2 rhs[SB][NJ];
    double lm[SB][C0], rp[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nb + q;
      if (nn < nd) {
        const double *fr = FV.ff + ((size_t)nn * ncls + cls) * 3;
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) rhs[q][jj] = xin[jj][nn * sin];
#pragma unroll
        for (int j = 0; j < C0; j++) lm[q][j] = fr[j];
        rp[q] = fr[2];
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nb + q;
      if (nn < nd) {
        const int nne = nn < C0 ? nn : C0;
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) {
          double2 v = rhs[q][jj];
#pragma unroll
          for (int j = 1; j <= C0; j++)
            if (j <= nne) v = sub2(v, axc(lm[q][j - 1], xh[jj][j - 1]));
#pragma unroll
          for (int t = C0 - 1; t >= 1; t--) xh[jj][t] = xh[jj][t - 1];
          xh[jj][0] = v;
          rhs[q][jj] = axc(rp[q], v);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nb + q < nd)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) xt[jj][(nb + q) * st] = rhs[q][jj];
  }
  double2 yh[NJ][KU];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
#pragma unroll
    for (int u = 0; u < KU; u++) yh[jj][u] = zero2();
  for (int nt = nd - 1; nt >= 0; nt -= SB) {
    double2 v0[SB][NJ];
    double um[SB][KU];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nt - q;
      if (nn >= 0) {
        const double *fr = FV.fb + ((size_t)nn * ncls + cls) * 2;
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) v0[q][jj] = xt[jj][nn * st];
#pragma unroll
        for (int u = 0; u < KU; u++) um[q][u] = fr[u];
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nt - q;
      if (nn >= 0) {
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) {
          double2 v = v0[q][jj];
          if (nn <= nd - 2) {
#pragma unroll
            for (int i = 1; i <= KU; i++)
              if (nn + i < nd) v = sub2(v, axc(um[q][i - 1], yh[jj][i - 1]));
          }
          v0[q][jj] = v;
#pragma unroll
          for (int u = KU - 1; u >= 1; u--) yh[jj][u] = yh[jj][u - 1];
          yh[jj][0] = v;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nt - q >= 0)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) x[jj][nt - q] = v0[q][jj];
  }
}


// x <- XM x on nn rows (tri-band, ascending, from the original values)
template <int SB>
__device__ void mulxm_column(double2 *x, const double *__restrict__ xt, int nn) {
  double2 prev = zero2();
  for (int cb = 0; cb < nn; cb += SB) {
    double2 cur[SB + 1];
    double r[SB][3];
#pragma unroll
    for (int q = 0; q <= SB; q++)
      if (cb + q < nn) cur[q] = x[cb + q];
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (cb + q < nn) {
        r[q][0] = xt[(size_t)(cb + q) * 3];
        r[q][1] = xt[(size_t)(cb + q) * 3 + 1];
        r[q][2] = xt[(size_t)(cb + q) * 3 + 2];
      }
    double2 o[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int c = cb + q;
      if (c < nn) {
        double2 t = axc(r[q][1], cur[q]);
        if (c >= 1) t = add2(t, axc(r[q][0], prev));
        if (c + 1 < nn) t = add2(t, axc(r[q][2], cur[q + 1]));
        prev = cur[q];
        o[q] = t;
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (cb + q < nn) x[cb + q] = o[q];
  }
}

__device__ void solve_setup(const KGrid &g, const double *bd, const double *bxm, const double *ak, int op, double alp, int mm,
                            int kidx, SolveCtx &s) {
  const int mv = mm * g.minc;
  s.g = g;
  s.bd = bd;
  s.bxm = bxm;
  s.ak = ak;
  s.op = op;
  s.mm = mm;
  s.k = kidx;
  s.np = s_nrchops(g.nrchop, mv) + 1;
  double akv = ak[mm + (size_t)g.ntchopdim * kidx];
  s.k2 = akv * akv;
  s.alp = (op == 2) ? alp : 0.0;
  s.kind = 0;
  if (mm == 0) {
    if (op == 0 && kidx == 0) s.kind = 1;
    if (op == 2) s.kind = 2;
  }
}

// One thread per (m, |k|): factorise the operator's band matrix.
__global__ void __launch_bounds__(64) k_band_factor(KGrid g, const double *bd, const double *bxm, const double *ak, int op,
                                                    double alp, double *F) {
  const long ncls = (long)g.ntchopdim * g.nxchop;
  const long cls = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (cls >= ncls) return;
  const int mm = (int)(cls % g.ntchopdim), kabs = (int)(cls / g.ntchopdim);
  const int nn = s_nrchops(g.nrchop, mm * g.minc);
  if (nn < 1 || mm >= g.ntchop) return;
  SolveCtx s;
  solve_setup(g, bd, bxm, ak, op, alp, mm, kabs, s);
  const FactorView FV = factor_view(F, g.nrchopdim, ncls);
  if (s.kind == 0) band_factor<5, 2>(s, nn, FV, ncls, cls);
  else if (s.kind == 1) band_factor<7, 3>(s, nn, FV, ncls, cls);
  else band_factor<7, 3>(s, nn - 1, FV, ncls, cls);
}

// IDEL2 / IDEL2LN / IHELM (src/mod_legops.f90:823-1354) on the cached
// factors.  k_band_apply_penta: every (m, k) system solved with the
// penta-diagonal factors (m >= 1, and m = 0 where the operator keeps the
// penta-diagonal form), one thread each.  k_band_apply_m0: the m = 0
// systems with their own forms (IDEL2's Dirichlet log row at k = 0, IHELM's
// (1-x)-factorised 7-band system at every k), with the right-hand-side and
// result fix-ups, one thread per k.  LN outputs of m = 0, k = 0 are set by
// the thread that owns that system.
__device__ __forceinline__ int solve_kind(int op, int mm, int k) {
  if (mm != 0) return 0;
  if (op == 0 && k == 0) return 1;
  if (op == 2) return 2;
  return 0;
}

struct BandJob {
  const double2 *b;  // right-hand side (may be a)
  double2 *a;
  double *ln_a;
  const double *ln_b;
  const double *F;
  double alp;
  double2 *t;  // scratch of FFF size: forward-pass results, system-interleaved
};
constexpr int MAX_BAND_JOBS = 3;
struct BandJobs {
  BandJob j[MAX_BAND_JOBS];
  int bil;                                 // right-hand sides row-interleaved (row n of system s at b[n NSYS + s])
  int n;                                   // jobs
  int ng;                                  // groups of jobs sharing a factor set
  int gcount[MAX_BAND_JOBS];               // jobs of group g
  int gjob[MAX_BAND_JOBS][MAX_BAND_JOBS];  // and their indices
};

// LN of the m = 0, k = 0 results, set by the thread that wrote their row 0.
template <int NJ>
__device__ void penta_ln(const KGrid &g, int op, int has_ln, const BandJobs &jobs, int grp, double2 *const (&x)[NJ]) {
#pragma unroll
  for (int jj = 0; jj < NJ; jj++) {
    double *ln_a = jobs.j[jobs.gjob[grp][jj]].ln_a;
    if (op == 0) ln_a[0] = has_ln ? x[jj][0].x * g.norm0 : 0.0;
    if (op == 1) {
      ln_a[0] = x[jj][0].x * g.norm0;
      x[jj][0] = zero2();
    }
  }
}

// The systems of one (m, k) for the NJ jobs of group grp (one factor set).
template <int NJ>
__device__ void apply_penta_n(const KGrid &g, int op, int has_ln, const BandJobs &jobs, int grp, long sys) {
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  if (sys >= nsys) return;
  const int mm = (int)(sys % g.ntchopdim), k = (int)(sys / g.ntchopdim);
  const int nn = s_nrchops(g.nrchop, mm * g.minc);
  if (nn < 1 || mm >= g.ntchop || solve_kind(op, mm, k) != 0) return;
  const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
  const size_t col = (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k);
  const double2 *xin[NJ];
  double2 *x[NJ], *xt[NJ];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++) {
    const BandJob &J = jobs.j[jobs.gjob[grp][jj]];
    xin[jj] = jobs.bil ? J.b + sys : J.b + col;
    x[jj] = J.a + col;
    xt[jj] = J.t + sys;
  }
  const FactorView FV = factor_view(jobs.j[jobs.gjob[grp][0]].F, g.nrchopdim, ncls);
  band_apply_n<NJ, 5, 2, (NJ == 1 ? 16 : 8)>(nn, xin, jobs.bil ? nsys : 1L, xt, x, nsys, FV, ncls,
                                             mm + (long)g.ntchopdim * kabs);
  if (mm == 0 && k == 0) penta_ln<NJ>(g, op, has_ln, jobs, grp, x);
}

// LDS-staged partitioned substitution.  A block (one wave) takes W = 64 / P
// consecutive systems of one group, P lanes each: lane j P + seg owns rows
// [seg PART_LS, (seg + 1) PART_LS) of system j.
//  - The systems' right-hand sides come into LDS with coalesced loads, each
//    lane's rows in a slot of PART_LS + 1 (the pad spreads the lanes' rows
//    over the banks); the factor rows of a lane's segment come into its
//    registers.
//  - Each pass is a second-order linear recurrence in the unscaled values
//    (forward: v_n = b_n - l1_n v_{n-1} - l2_n v_{n-2}; back: x_n = y_n -
//    u1_n x_{n+1} - u2_n x_{n+2}), so a segment's end state is E = p + G s
//    for its start state s: p from a run over the segment's rows from the
//    zero state, G's columns from runs from the unit states without right-
//    hand side.  The start states pass from lane to lane (part_carry: P - 1
//    steps of DPP moves), then every lane runs its rows again from its own
//    start state.  The forward results overwrite the right-hand side in
//    LDS, the back substitution's results overwrite those, and the block
//    stores the columns with coalesced writes.
//  - Every lane runs all PART_LS rows with no branch: rows at or beyond the
//    system's end, and the couplings LUBKSBB skips (rows 0, 1 forward, the
//    last two rows back), take zero factors and right-hand sides, which
//    carry a zero state through exactly.
//  - The first two segments of each pass are LUBKSBB's arithmetic exactly
//    (their start states are exact); the others differ from it by the
//    rounding of the carried states.
#define PART_LS 8
__host__ __device__ constexpr size_t part_lds_bytes(int nj) { return sizeof(double2) * nj * 64 * (PART_LS + 1); }

// Rows of penta-diagonal system sys under operator op (0: not one, or no
// rows), with its m index and k slot.
__device__ __forceinline__ int penta_rows(const KGrid &g, int op, long sys, int &mm, int &k) {
  mm = k = 0;
  if (sys >= (long)g.ntchopdim * g.nxchopdim) return 0;
  const int m = (int)(sys % g.ntchopdim), kk = (int)(sys / g.ntchopdim);
  const int nn = s_nrchops(g.nrchop, m * g.minc);
  if (nn < 1 || m >= g.ntchop || solve_kind(op, m, kk) != 0) return 0;
  mm = m;
  k = kk;
  return nn;
}

__device__ __forceinline__ double2 sel2(bool c, double2 v) { return c ? v : zero2(); }
// Compiler-only fence: the rows of a block are read from LDS where the block
// starts, not kept in registers from an earlier pass or hoisted ahead.
__device__ __forceinline__ void lds_fence() { asm volatile("" ::: "memory"); }

// One row of a pass with C couplings: v = r - c_0 h_0 - ... - c_{C-1} h_{C-1},
// h_t the value t + 1 rows back in the pass's direction, then the state
// shifts.
template <int NJ, int C>
__device__ __forceinline__ void part_row(const double (&c)[C], const double2 (&r)[NJ], double2 (&h)[NJ][C]) {
#pragma unroll
  for (int jj = 0; jj < NJ; jj++) {
    double2 v = r[jj];
#pragma unroll
    for (int t = 0; t < C; t++) v = sub2(v, axc(c[t], h[jj][t]));
#pragma unroll
    for (int t = C - 1; t >= 1; t--) h[jj][t] = h[jj][t - 1];
    h[jj][0] = v;
  }
}
// The same row on a unit-state run (no right-hand side, real)
template <int C>
__device__ __forceinline__ void part_unit(const double (&c)[C], double (&g)[C]) {
  double t = 0.0;
#pragma unroll
  for (int u = 0; u < C; u++) t = t - c[u] * g[u];
#pragma unroll
  for (int u = C - 1; u >= 1; u--) g[u] = g[u - 1];
  g[0] = t;
}

// Values of the lane below (dir = 1) or above (dir = -1) in the wave, moved
// by DPP (wave_shr:1 / wave_shl:1); lane 0, resp. 63, gets zero.
template <int DIR>
__device__ __forceinline__ double dpp_next(double v) {
  constexpr int ctrl = DIR > 0 ? 0x138 : 0x130;
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffff), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int DIR>
__device__ __forceinline__ double2 dpp_next2(double2 v) {
  return make_double2(dpp_next<DIR>(v.x), dpp_next<DIR>(v.y));
}

// The start states of the segments, passed along the P lanes of each system
// in turn (lanes seg + 1 after seg forward, seg - 1 after seg back): every
// lane forms its segment's end state E = p + G s from its current start
// state (G[u]: the end state of the run from unit state u), and at step t
// the lane of segment t (forward) or P - 1 - t (back) takes its
// neighbour's.  Each carried value is a state of the recurrence, as in the
// serial substitution.
template <int NJ, int C, int DIR>
__device__ __forceinline__ void part_carry(int P, int seg, const double2 (&p)[NJ][C], const double (&G)[C][C],
                                           double2 (&st)[NJ][C]) {
  for (int t = 1; t < P; t++) {
    const int take = DIR > 0 ? t : P - 1 - t;
#pragma unroll
    for (int jj = 0; jj < NJ; jj++) {
      double2 r[C];
#pragma unroll
      for (int q = 0; q < C; q++) {
        double2 acc = axc(G[0][q], st[jj][0]);
#pragma unroll
        for (int u = 1; u < C; u++) acc = add2(acc, axc(G[u][q], st[jj][u]));
        r[q] = dpp_next2<DIR>(add2(p[jj][q], acc));
      }
      if (seg == take) {
#pragma unroll
        for (int q = 0; q < C; q++) st[jj][q] = r[q];
      }
    }
  }
}

// The passes of the partitioned substitution on rows staged in LDS (my:
// this lane's slot of PART_LS + 1 rows; field jj at my + jj fstride), for a
// lane owning rows a .. a + PART_LS - 1 of a system of nd rows, C couplings
// per row.  fwd(n, c, sc) gives the couplings of row n of the forward pass
// (c[t] to the value of row n - 1 - t) and the scale of its result,
// back(n, c) those of the back substitution (c[t] to row n + 1 + t); both
// give zero couplings where the serial substitution has none.  The back
// pass's couplings are loaded after the forward pass, when the forward ones
// are dead.
template <int NJ, int C, class Fwd, class Back>
__device__ __forceinline__ void part_passes(double2 *my, size_t fstride, int a, int nd, int P, int seg, Fwd fwd,
                                            Back back) {
  constexpr int LS = PART_LS, SB = 4;
  double cf[LS][C], rp[LS];
#pragma unroll
  for (int q = 0; q < LS; q++) fwd(a + q, cf[q], rp[q]);
  double2 st[NJ][C], p[NJ][C];
  // ---- forward pass (rows ascending)
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
#pragma unroll
    for (int t = 0; t < C; t++) st[jj][t] = p[jj][t] = zero2();
  {
    double G[C][C];
#pragma unroll
    for (int u = 0; u < C; u++)
#pragma unroll
      for (int t = 0; t < C; t++) G[u][t] = u == t ? 1.0 : 0.0;
#pragma unroll
    for (int qb = 0; qb < LS; qb += SB) {
      double2 r[SB][NJ];
      lds_fence();
#pragma unroll
      for (int q = 0; q < SB; q++)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) r[q][jj] = sel2(a + qb + q < nd, my[jj * fstride + qb + q]);
#pragma unroll
      for (int q = 0; q < SB; q++) {
#pragma unroll
        for (int u = 0; u < C; u++) part_unit<C>(cf[qb + q], G[u]);
        part_row<NJ, C>(cf[qb + q], r[q], p);
      }
    }
    part_carry<NJ, C, 1>(P, seg, p, G, st);
  }
#pragma unroll
  for (int qb = 0; qb < LS; qb += SB) {
    double2 r[SB][NJ];
    lds_fence();
#pragma unroll
    for (int q = 0; q < SB; q++)
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) r[q][jj] = sel2(a + qb + q < nd, my[jj * fstride + qb + q]);
#pragma unroll
    for (int q = 0; q < SB; q++) {
      part_row<NJ, C>(cf[qb + q], r[q], st);
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) my[jj * fstride + qb + q] = axc(rp[qb + q], st[jj][0]);
    }
  }
  // ---- back substitution (rows descending)
  double cb[LS][C];
#pragma unroll
  for (int q = 0; q < LS; q++) back(a + q, cb[q]);
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
#pragma unroll
    for (int t = 0; t < C; t++) st[jj][t] = p[jj][t] = zero2();
  {
    double G[C][C];
#pragma unroll
    for (int u = 0; u < C; u++)
#pragma unroll
      for (int t = 0; t < C; t++) G[u][t] = u == t ? 1.0 : 0.0;
#pragma unroll
    for (int qb = LS - SB; qb >= 0; qb -= SB) {
      double2 r[SB][NJ];
      lds_fence();
#pragma unroll
      for (int q = 0; q < SB; q++)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) r[q][jj] = sel2(a + qb + q < nd, my[jj * fstride + qb + q]);
#pragma unroll
      for (int q = SB - 1; q >= 0; q--) {
#pragma unroll
        for (int u = 0; u < C; u++) part_unit<C>(cb[qb + q], G[u]);
        part_row<NJ, C>(cb[qb + q], r[q], p);
      }
    }
    part_carry<NJ, C, -1>(P, seg, p, G, st);
  }
#pragma unroll
  for (int qb = LS - SB; qb >= 0; qb -= SB) {
    double2 r[SB][NJ];
    lds_fence();
#pragma unroll
    for (int q = 0; q < SB; q++)
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) r[q][jj] = sel2(a + qb + q < nd, my[jj * fstride + qb + q]);
#pragma unroll
    for (int q = SB - 1; q >= 0; q--) {
      part_row<NJ, C>(cb[qb + q], r[q], st);
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) my[jj * fstride + qb + q] = st[jj][0];
    }
  }
}

template <int NJ>
__device__ void apply_penta_part(const KGrid &g, int op, int has_ln, const BandJobs &jobs, int grp, long sys0, int P,
                                 double2 *lds) {
  constexpr int LS = PART_LS, SLOT = PART_LS + 1;
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  const int lane = threadIdx.x, W = 64 / P;
  const int j = lane / P, seg = lane % P;
  int mm, k;
  const int nd = penta_rows(g, op, sys0 + j, mm, k);
  const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
  const long cls = mm + (long)g.ntchopdim * kabs;
  const FactorView FV = factor_view(jobs.j[jobs.gjob[grp][0]].F, g.nrchopdim, ncls);
  const int nlast = max(nd - 1, 0);
  // right-hand sides into LDS: system jp's row n at slot (jp P + n / LS), row n % LS
  const size_t fstride = (size_t)64 * SLOT;
  if (jobs.bil) {
    // row n of system s at b[n nsys + s]: lanes over W consecutive systems
    int m2, k2;
    const int jp = lane % W;
    const int ndp = penta_rows(g, op, sys0 + jp, m2, k2);
    for (int n = lane / W; n < ndp; n += P) {
      const int sl = (jp * P + n / LS) * SLOT + n % LS;
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) lds[jj * fstride + sl] = jobs.j[jobs.gjob[grp][jj]].b[(size_t)n * nsys + sys0 + jp];
    }
  } else {
    for (int jp = 0; jp < W; jp++) {
      int m2, k2;
      const int ndp = penta_rows(g, op, sys0 + jp, m2, k2);
      const size_t colp = (size_t)g.nrchopdim * ((size_t)m2 + (size_t)g.ntchopdim * k2);
      for (int n = lane; n < ndp; n += 64) {
        const int sl = (jp * P + n / LS) * SLOT + n % LS;
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) lds[jj * fstride + sl] = jobs.j[jobs.gjob[grp][jj]].b[colp + n];
      }
    }
  }
  __syncthreads();
  double2 *my = lds + (size_t)lane * SLOT;
  // LUBKSBB: row n's multipliers against rows n-1, n-2 and 1/pivot, then its
  // normalised upper entries against rows n+1, n+2
  part_passes<NJ, 2>(
      my, fstride, seg * LS, nd, P, seg,
      [&](int n, double(&c)[2], double &sc) {
        const double *ff = FV.ff + ((size_t)min(n, nlast) * ncls + cls) * 3;
        c[0] = (n >= 1 && n < nd) ? ff[0] : 0.0;
        c[1] = (n >= 2 && n < nd) ? ff[1] : 0.0;
        sc = ff[2];
      },
      [&](int n, double(&c)[2]) {
        const double *fb = FV.fb + ((size_t)min(n, nlast) * ncls + cls) * 2;
        c[0] = n + 1 < nd ? fb[0] : 0.0;
        c[1] = n + 2 < nd ? fb[1] : 0.0;
      });
  __syncthreads();
  // LN of the m = 0, k = 0 results (row 0 of segment 0), then the columns out
  if (seg == 0 && nd > 0 && mm == 0 && k == 0) {
#pragma unroll
    for (int jj = 0; jj < NJ; jj++) {
      double *ln_a = jobs.j[jobs.gjob[grp][jj]].ln_a;
      if (op == 0) ln_a[0] = has_ln ? my[jj * fstride].x * g.norm0 : 0.0;
      if (op == 1) {
        ln_a[0] = my[jj * fstride].x * g.norm0;
        my[jj * fstride] = zero2();
      }
    }
  }
  __syncthreads();
  for (int jp = 0; jp < W; jp++) {
    int m2, k2;
    const int ndp = penta_rows(g, op, sys0 + jp, m2, k2);
    const size_t colp = (size_t)g.nrchopdim * ((size_t)m2 + (size_t)g.ntchopdim * k2);
    for (int n = lane; n < ndp; n += 64) {
      const int sl = (jp * P + n / LS) * SLOT + n % LS;
#pragma unroll
      for (int jj = 0; jj < NJ; jj++) jobs.j[jobs.gjob[grp][jj]].a[colp + n] = lds[jj * fstride + sl];
    }
  }
}

// The m = 0 systems (one per k) are the longest dependency chains of a
// solve.  One 64-thread block per system stages the system's factor rows,
// right-hand side and XM rows in LDS with coalesced loads; lane 0 then runs
// the substitutions from LDS (the same arithmetic as from memory) and the
// block writes the column back.
#define M0_SB 4
size_t m0_lds_bytes(const KGrid &g) { return sizeof(double) * 12 * (size_t)g.nrchopdim; }

__device__ void apply_m0(const KGrid &g, const double *bxm, int op, double alp, int has_ln, double lnval,
                         const double2 *b, double2 *a, double *ln_a, const double *ln_b, const double *__restrict__ F,
                         int k, double *sm, int bil) {
  const long ncls = (long)g.ntchopdim * g.nxchop;
  if (k >= g.nxchopdim) return;
  // IHELM: A%LN starts as B%LN (COPY), replaced below by -B%LN/ALP
  if (op == 2 && k == 0 && threadIdx.x == 0) ln_a[0] = ln_b[0] * 1.0 + 0.0;
  const int nn = s_nrchops(g.nrchop, 0);
  const int kind = solve_kind(op, 0, k);
  if (nn < 1 || g.ntchop < 1 || kind == 0) return;
  const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
  const long cls = (long)g.ntchopdim * kabs;
  // the right-hand side's rows: column (m = 0, k), or row-interleaved (system NTCHOPDIM k)
  const size_t bstride = bil ? (size_t)g.ntchopdim * g.nxchopdim : 1;
  const double2 *xb = bil ? b + (size_t)g.ntchopdim * k : b + (size_t)g.nrchopdim * ((size_t)g.ntchopdim * k);
  double2 *xg = a + (size_t)g.nrchopdim * ((size_t)g.ntchopdim * k);
  double2 *x = reinterpret_cast<double2 *>(sm);  // [nrchopdim]
  double *Fs = sm + 2 * g.nrchopdim;             // [nrchopdim][FSTRIDE]
  double *xt = Fs + FSTRIDE * g.nrchopdim;       // [nrchopdim][3]
  const int nd = kind == 1 ? nn : nn - 1;        // factorised rows
  // IDEL2 m0k0: the first row is the log-term (Dirichlet) row, the right-
  // hand side moves down one row under it
  if (kind == 1 && threadIdx.x == 0) x[0] = make_double2(has_ln ? lnval / g.norm0 : 0.0, 0.0);
  for (int row = threadIdx.x; row < nn; row += blockDim.x) {
    if (kind == 2) x[row] = xb[row * bstride];
    else if (row + 1 < nn) x[row + 1] = xb[row * bstride];
    if (row < nd) {
      const double *fr = factor_view(F, g.nrchopdim, ncls).f7 + ((size_t)row * ncls + cls) * FSTRIDE;
#pragma unroll
      for (int e = 0; e < FSTRIDE; e++) Fs[row * FSTRIDE + e] = fr[e];
    }
    if (kind == 2) {
#pragma unroll
      for (int e = 0; e < 3; e++) xt[row * 3 + e] = bxm[(size_t)row * 3 + e];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (kind == 1) {
      band_apply<7, 3, M0_SB>(nn, x, x, Fs, 1, 0);
      ln_a[0] = has_ln ? x[0].x * g.norm0 : 0.0;
    } else {
      // IHELM m=0: solve (del^2 - alp - k^2)(1-x) y = rhs on NN-1 rows, A = (1-x) y
      if (k == 0) {
        double lnA = -ln_b[0] / alp;
        ln_a[0] = lnA;
        x[0].x = x[0].x - lnA * (4.0 / 3 / g.ell2) / g.norm0;
        x[1].x = x[1].x - lnA * (-2 / g.norm1 / g.ell2);
        x[2].x = x[2].x - lnA * (+2.0 / 3 / g.norm2 / g.ell2);
      }
      band_apply<7, 3, M0_SB>(nn - 1, x, x, Fs, 1, 0);
      x[nn - 1] = zero2();
      mulxm_column<M0_SB>(x, xt, nn);
    }
  }
  __syncthreads();
  for (int row = threadIdx.x; row < nn; row += blockDim.x) xg[row] = x[row];
}

// One launch: blocks [0, nm0) take the m = 0 systems (one block each, first
// because they are the longest), the remaining blocks the penta-diagonal
// systems, so both sets run side by side; grid row y solves field y of a
// batch of independent solves.  k_band_apply: one thread per penta-diagonal
// system; k_band_apply_part: P lanes per system, 64 / P systems per block
// (apply_penta_part), at three waves per SIMD.
__device__ __forceinline__ bool band_m0_block(const KGrid &g, const double *bxm, int op, int has_ln, double lnval,
                                              const BandJobs &jobs, int nm0, double *m0s) {
  const int bid = blockIdx.x;
  if (bid >= nm0 * jobs.n) return false;
  const BandJob &J = jobs.j[bid / nm0];
  apply_m0(g, bxm, op, J.alp, has_ln, lnval, J.b, J.a, J.ln_a, J.ln_b, J.F, bid % nm0, m0s, jobs.bil);
  return true;
}

__global__ void __launch_bounds__(64) k_band_apply(KGrid g, const double *bxm, int op, int has_ln, double lnval, BandJobs jobs,
                                                   int nm0, int npenta) {
  extern __shared__ __align__(16) double m0s[];
  if (band_m0_block(g, bxm, op, has_ln, lnval, jobs, nm0, m0s)) return;
  const int pb = blockIdx.x - nm0 * jobs.n, grp = pb / npenta;
  const long sys = (long)(pb - grp * npenta) * blockDim.x + threadIdx.x;
  switch (jobs.gcount[grp]) {
    case 1: apply_penta_n<1>(g, op, has_ln, jobs, grp, sys); break;
    case 2: apply_penta_n<2>(g, op, has_ln, jobs, grp, sys); break;
    default: apply_penta_n<3>(g, op, has_ln, jobs, grp, sys); break;
  }
}

// The m = 0 systems of apply_m0 on the partitioned passes, all 64 lanes of
// the block on one system (lane l owns rows 8 l .. 8 l + 7; NRCHOP <= 512):
// IHELM's (1-x)-factorised 7-band system on NN-1 rows followed by XM, with
// its LN terms; three couplings per row.  IDEL2's Dirichlet log-row system
// (k = 0) stays on apply_m0's serial substitution: carried states lose up to
// four digits on it (its homogeneous solutions grow fast), against 1e-14 on
// IHELM's.
__device__ void apply_m0_part(const KGrid &g, const double *bxm, int op, double alp, int has_ln, double lnval,
                              const double2 *b, double2 *a, double *ln_a, const double *ln_b, const double *__restrict__ F,
                              int k, double2 *lds, int bil) {
  if (solve_kind(op, 0, k) == 1) {
    apply_m0(g, bxm, op, alp, has_ln, lnval, b, a, ln_a, ln_b, F, k, reinterpret_cast<double *>(lds), bil);
    return;
  }
  constexpr int LS = PART_LS, SLOT = PART_LS + 1;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  const int lane = threadIdx.x;
  if (k >= g.nxchopdim) return;
  // IHELM: A%LN starts as B%LN (COPY), replaced below by -B%LN/ALP
  if (op == 2 && k == 0 && lane == 0) ln_a[0] = ln_b[0] * 1.0 + 0.0;
  const int nn = s_nrchops(g.nrchop, 0);
  const int kind = solve_kind(op, 0, k);
  if (nn < 1 || g.ntchop < 1 || kind == 0) return;  // the whole block
  const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
  const long cls = (long)g.ntchopdim * kabs;
  const size_t bstride = bil ? (size_t)g.ntchopdim * g.nxchopdim : 1;
  const double2 *xb = bil ? b + (size_t)g.ntchopdim * k : b + (size_t)g.nrchopdim * ((size_t)g.ntchopdim * k);
  double2 *xg = a + (size_t)g.nrchopdim * ((size_t)g.ntchopdim * k);
  auto slot = [&](int n) { return (n / LS) * SLOT + n % LS; };
  const double lnA = k == 0 ? -ln_b[0] / alp : 0.0;
  for (int n = lane; n < nn; n += 64) {
    double2 v = xb[n * bstride];
    if (k == 0) {
      if (n == 0) v.x = v.x - lnA * (4.0 / 3 / g.ell2) / g.norm0;
      if (n == 1) v.x = v.x - lnA * (-2 / g.norm1 / g.ell2);
      if (n == 2) v.x = v.x - lnA * (+2.0 / 3 / g.norm2 / g.ell2);
    }
    lds[slot(n)] = v;
  }
  if (k == 0 && lane == 0) ln_a[0] = lnA;
  __syncthreads();
  const int nd = nn - 1;  // factorised rows
  const double *f7 = factor_view(F, g.nrchopdim, ncls).f7;
  const int nlast = max(nd - 1, 0);
  auto rec = [&](int n) { return f7 + ((size_t)min(max(n, 0), nlast) * ncls + cls) * FSTRIDE; };
  part_passes<1, 3>(
      lds + (size_t)lane * SLOT, 0, lane * LS, nd, 64, lane,
      [&](int n, double(&c)[3], double &sc) {
        const double *fr = rec(n);
#pragma unroll
        for (int t = 0; t < 3; t++) c[t] = (n >= t + 1 && n < nd) ? fr[t] : 0.0;
        sc = fr[6];
      },
      [&](int n, double(&c)[3]) {
        const double *fr = rec(n);
#pragma unroll
        for (int t = 0; t < 3; t++) c[t] = n + t + 1 < nd ? fr[3 + t] : 0.0;
      });
  __syncthreads();
  // the (1-x) y solve ends with row NN-1 zero, then A = XM y on NN
  // rows (tri-band, from the solved values, ascending sums as mulxm_column)
  double2 o[LS];
#pragma unroll
  for (int q = 0; q < LS; q++) {
    const int c = lane * LS + q;
    o[q] = zero2();
    if (c < nn) {
      const double2 cur = c < nn - 1 ? lds[slot(c)] : zero2();
      double2 t = axc(bxm[(size_t)c * 3 + 1], cur);
      if (c >= 1) t = add2(t, axc(bxm[(size_t)c * 3 + 0], lds[slot(c - 1)]));
      if (c + 1 < nn) t = add2(t, axc(bxm[(size_t)c * 3 + 2], c + 1 < nn - 1 ? lds[slot(c + 1)] : zero2()));
      o[q] = t;
    }
  }
#pragma unroll
  for (int q = 0; q < LS; q++) {
    const int c = lane * LS + q;
    if (c < nn) xg[c] = o[q];
  }
}

__global__ void __launch_bounds__(64, 3) k_band_apply_part(KGrid g, const double *bxm, int op, int has_ln, double lnval,
                                                           BandJobs jobs, int nm0, int npenta, int P) {
  extern __shared__ __align__(16) double m0s[];
  if ((int)blockIdx.x < nm0 * jobs.n) {
    const BandJob &J = jobs.j[blockIdx.x / nm0];
    apply_m0_part(g, bxm, op, J.alp, has_ln, lnval, J.b, J.a, J.ln_a, J.ln_b, J.F, blockIdx.x % nm0,
                  reinterpret_cast<double2 *>(m0s), jobs.bil);
    return;
  }
  // the penta-diagonal blocks from the next multiple of 8 on, XCD-aware: the
  // blocks of one XCD take consecutive systems and share factor-row lines
  const int pb = xcd_logical(npenta * jobs.ng, xcd_grid(nm0 * jobs.n));
  if (pb < 0) return;
  const int grp = pb / npenta;
  const long sys0 = (long)(pb - grp * npenta) * (64 / P);
  double2 *lds = reinterpret_cast<double2 *>(m0s);
  switch (jobs.gcount[grp]) {
    case 1: apply_penta_part<1>(g, op, has_ln, jobs, grp, sys0, P, lds); break;
    case 2: apply_penta_part<2>(g, op, has_ln, jobs, grp, sys0, P, lds); break;
    default: apply_penta_part<3>(g, op, has_ln, jobs, grp, sys0, P, lds); break;
  }
}

// Lanes per penta-diagonal system of the partitioned solve: the smallest
// power of two, at least 4, whose segments of PART_LS rows cover NRCHOP.
// 0: one thread per system (NRCHOP below 64, where systems are short, or
// above 64 PART_LS); that path is LUBKSBB's arithmetic, bitwise.
int band_lanes(const KGrid &g) {
  if (g.nrchop < 64 || g.nrchop > 64 * PART_LS) return 0;
  int p = 4;
  while (p * PART_LS < g.nrchop) p *= 2;
  return p;
}

// Factor sets are cached per (operator, alp, truncation): the matrices
// depend on m, k^2 and those only, so repeated solves (every CN step) only
// substitute.  Results are bitwise those of factorising on every call.
// At most eight sets are kept; a miss evicts the oldest set that neither
// the current batch nor a captured step uses (pinned sets are never
// evicted; with all sets pinned the cache grows past eight).
static double *factor_set(mlg_ctx *c, int op, double alp, const double *const *keep, int nkeep, int *rc) {
  KGrid g = kgrid(c);
  double key_alp = (op == 2) ? alp : 0.0;
  for (auto &f : c->fcache)
    if (f.op == op && f.alp == key_alp && f.nrchop == g.nrchop) return f.f;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  size_t n = (size_t)g.nrchopdim * FSET_DOUBLES * ncls;
  if (c->fcache.size() >= 8) {
    for (auto it = c->fcache.begin(); it != c->fcache.end(); ++it) {
      bool used = false;
      for (int i = 0; i < nkeep; i++) used = used || keep[i] == it->f;
      if (used || it->pins > 0) continue;
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(it->f);
      c->fcache.erase(it);
      break;
    }
  }
  double *F = nullptr;
  if (hipMalloc((void **)&F, sizeof(double) * n) != hipSuccess) {
    *rc = MLG_E_NOMEM;
    return nullptr;
  }
  hipLaunchKernelGGL(k_band_factor, dim3((ncls + 63) / 64), dim3(64), 0, c->stream, g, c->d.bdel2h, c->d.bxm, c->d.ak, op,
                     key_alp, F);
  if (hipGetLastError() != hipSuccess) {
    (void)hipFree(F);
    *rc = MLG_E_HIP;
    return nullptr;
  }
  c->fcache.push_back(FactorSet{op, key_alp, g.nrchop, F});
  return F;
}

int launch_band_solve_n(mlg_ctx *c, int op, int n, const double2 *const *b, double2 *const *a, double *const *ln_a,
                        const double *const *ln_b, const double *alp, int has_ln, double ln, double2 *const *t, int bil) {
  if (n < 1 || n > MAX_BAND_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  long nsys = (long)g.ntchopdim * g.nxchopdim;
  BandJobs jobs;
  const double *used[MAX_BAND_JOBS];
  for (int i = 0; i < n; i++) {
    int rc = 0;
    const double *F = factor_set(c, op, alp[i], used, i, &rc);
    if (!F) return rc;
    used[i] = F;
    if (!t[i]) return MLG_E_ARG;
    jobs.j[i] = BandJob{b[i], a[i], ln_a[i], ln_b[i], F, alp[i], t[i]};
  }
  jobs.n = n;
  jobs.bil = bil;
  jobs.ng = 0;
  for (int i = 0; i < n; i++) {
    int gi = 0;
    while (gi < jobs.ng && jobs.j[jobs.gjob[gi][0]].F != jobs.j[i].F) gi++;
    if (gi == jobs.ng) jobs.gcount[jobs.ng++] = 0;
    jobs.gjob[gi][jobs.gcount[gi]++] = i;
  }
  KScope ks(c, K_BAND_SOLVE, n * 2.0 * 16.0 * fff_count(g));  // the coefficients read and written once
  const int lanes = band_lanes(g);
  const int nm0 = (op != 1) ? g.nxchopdim : 0;
  size_t lds = nm0 ? m0_lds_bytes(g) : 0;
  if (lanes) {
    int njmax = 1;
    for (int i = 0; i < jobs.ng; i++) njmax = std::max(njmax, jobs.gcount[i]);
    // the partitioned m = 0 blocks use one field's; IDEL2's serial one its own
    lds = std::max(part_lds_bytes(njmax), op == 0 ? m0_lds_bytes(g) : (size_t)0);
    const int npenta = (int)((nsys + 64 / lanes - 1) / (64 / lanes));
    hipLaunchKernelGGL(k_band_apply_part, dim3(xcd_grid(nm0 * n) + xcd_grid(npenta * jobs.ng)), dim3(64), lds, c->stream,
                       g, c->d.bxm, op, has_ln, ln, jobs, nm0, npenta, lanes);
  } else {
    const int npenta = (int)((nsys + 63) / 64);
    hipLaunchKernelGGL(k_band_apply, dim3(nm0 * n + npenta * jobs.ng), dim3(64), lds, c->stream, g, c->d.bxm, op, has_ln,
                       ln, jobs, nm0, npenta);
  }
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_band_solve(mlg_ctx *c, int op, const double2 *b, double2 *a, double *ln_a, const double *ln_b, double alp,
                      int has_ln, double ln, double2 *t) {
  return launch_band_solve_n(c, op, 1, &b, &a, &ln_a, &ln_b, &alp, has_ln, ln, &t, 0);
}

// ---------------------------------------------------------------- transposed solves
// Adjoint of IDEL2LN / IHELM: G = A^-T H for every system, on the factors
// the forward solve caches.  A = L D U with L, U unit triangular (L the
// stored multipliers, U the stored normalised upper entries, D the pivots),
// so A^T = U^T D L^T: U^T w = h (forward over the rows), v = D^-1 w, then
// L^T g = v (backward over the rows).  One thread per system, in place on
// the FFF column.
template <int NE, int C0, int SB>
__device__ void band_apply_t(int nd, double2 *x, const FactorView &FV, long ncls, long cls) {
  constexpr int KU = NE - 1 - C0;
  auto lmul = [&](int row, int j) -> double {  // L(row, row - j)
    if (NE == 5) return FV.ff[((size_t)row * ncls + cls) * 3 + (j - 1)];
    return FV.f7[((size_t)row * ncls + cls) * FSTRIDE + (j - 1)];
  };
  auto uup = [&](int row, int u) -> double {  // U(row, row + u)
    if (NE == 5) return FV.fb[((size_t)row * ncls + cls) * 2 + (u - 1)];
    return FV.f7[((size_t)row * ncls + cls) * FSTRIDE + C0 + (u - 1)];
  };
  auto rpiv = [&](int row) -> double {
    if (NE == 5) return FV.ff[((size_t)row * ncls + cls) * 3 + 2];
    return FV.f7[((size_t)row * ncls + cls) * FSTRIDE + NE - 1];
  };
  // rows in blocks of SB: a block's factor entries are loaded before its
  // serial arithmetic (out-of-range entries as 0, which leaves the sums exact)
  double2 wh[KU];
#pragma unroll
  for (int u = 0; u < KU; u++) wh[u] = zero2();
  for (int nb = 0; nb < nd; nb += SB) {
    double um[SB][KU], rp[SB];
    double2 h[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nb + q;
      if (nn < nd) {
        h[q] = x[nn];
        rp[q] = rpiv(nn);
#pragma unroll
        for (int u = 1; u <= KU; u++) um[q][u - 1] = nn - u >= 0 ? uup(nn - u, u) : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      if (nb + q < nd) {
        double2 w = h[q];
#pragma unroll
        for (int u = 1; u <= KU; u++) w = sub2(w, axc(um[q][u - 1], wh[u - 1]));
#pragma unroll
        for (int u = KU - 1; u >= 1; u--) wh[u] = wh[u - 1];
        wh[0] = w;
        h[q] = axc(rp[q], w);
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nb + q < nd) x[nb + q] = h[q];
  }
  double2 gh[C0];
#pragma unroll
  for (int j = 0; j < C0; j++) gh[j] = zero2();
  for (int nt = nd - 1; nt >= 0; nt -= SB) {
    double lm[SB][C0];
    double2 v[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int nn = nt - q;
      if (nn >= 0) {
        v[q] = x[nn];
#pragma unroll
        for (int j = 1; j <= C0; j++) lm[q][j - 1] = nn + j < nd ? lmul(nn + j, j) : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      if (nt - q >= 0) {
        double2 g = v[q];
#pragma unroll
        for (int j = 1; j <= C0; j++) g = sub2(g, axc(lm[q][j - 1], gh[j - 1]));
#pragma unroll
        for (int j = C0 - 1; j >= 1; j--) gh[j] = gh[j - 1];
        gh[0] = g;
        v[q] = g;
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (nt - q >= 0) x[nt - q] = v[q];
  }
}

// One system of the transposed solve, on its column x in LDS (rows >= NN
// already zero).  op 1 (IDEL2LN): the m = 0 columns first take the
// transpose of the fix at infinity (x(n) -= AT1(n)/AT1(0) x(0)); the
// (m = 0, k = 0) system's first unknown became LN (LN = x(0) NORM(0,0),
// x(0) = 0), so its adjoint enters as x(0) = NORM(0,0) LNbar.  op 2 (IHELM):
// the m = 0 systems solve (del^2 - alp - k^2)(1-x) y on NN-1 rows and
// return XM y, so XM^T comes first; at k = 0 the right-hand side had LN
// terms removed (A%LN = -B%LN/ALP), whose adjoint gives B's LNbar.  ln:
// LNbar of the solution; ln_out: of the right-hand side (op 2).
__device__ void solve_t_system(const KGrid &g, const double *bxm, const double *at1, int op, double alp, const FactorView &FV,
                               long ncls, int mm, int k, int nn, double2 *col, const double *ln, double *ln_out) {
  const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
  const long cls = mm + (long)g.ntchopdim * kabs;
  const int kind = solve_kind(op, mm, k);
  if (kind == 0) {
    if (op == 1 && mm == 0) {
      const double2 a0 = col[0];
      for (int n = 0; n < g.nrchop; n++) col[n] = sub2(col[n], axc(at1[n] / at1[0], a0));
      if (k == 0) col[0] = make_double2(g.norm0 * ln[0], 0.0);
    }
    band_apply_t<5, 2, 8>(nn, col, FV, ncls, cls);
    return;
  }
  // IHELM, m = 0: x <- XM^T x on NN rows (ascending, from the original values)
  double2 prev = zero2();
  for (int c = 0; c < nn; c++) {
    const double *r = bxm + (size_t)c * 3;
    double2 t = axc(r[1], col[c]);
    if (c + 1 < nn) t = add2(t, axc(bxm[(size_t)(c + 1) * 3 + 0], col[c + 1]));
    if (c >= 1) t = add2(t, axc(bxm[(size_t)(c - 1) * 3 + 2], prev));
    prev = col[c];
    col[c] = t;
  }
  band_apply_t<7, 3, 8>(nn - 1, col, FV, ncls, cls);
  col[nn - 1] = zero2();
  if (k == 0) {
    const double lbar = ln[0] - (col[0].x * ((4.0 / 3 / g.ell2) / g.norm0) + col[1].x * (-2 / g.norm1 / g.ell2) +
                                 col[2].x * (+2.0 / 3 / g.norm2 / g.ell2));
    ln_out[0] = -lbar / alp;
  }
}

struct BandTJobs {
  BandTJob j[MAX_BAND_JOBS];
};
// Block (group of spb consecutive systems, job): the group's columns (one
// contiguous run of the FFF array) are staged in LDS with coalesced loads,
// chopped to NRCHOPS(m) on the way in; one thread per system substitutes
// from LDS; the block writes the columns back, zeros beyond NRCHOPS(m)
// included (the solve's CHOPDO).  in may equal out.
__global__ void __launch_bounds__(64) k_band_apply_t(KGrid g, const double *bxm, const double *at1, int op, BandTJobs jobs,
                                                     int spb) {
  extern __shared__ __align__(16) double2 xs[];
  const BandTJob &J = jobs.j[blockIdx.y];
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  const int R = g.nrchopdim;
  const long sys0 = blockIdx.x * (long)spb;
  const size_t base = (size_t)R * sys0;  // column of system s = m + NTCHOPDIM k starts at R s
  const long nv = (long)spb * R;
  for (long i = threadIdx.x; i < nv; i += blockDim.x) {
    const long sys = sys0 + i / R;
    const int row = (int)(i % R);
    double2 v = zero2();
    if (sys < nsys) {
      const int mm = (int)(sys % g.ntchopdim);
      const int nn = mm < g.ntchop ? s_nrchops(g.nrchop, mm * g.minc) : 0;
      if (row < nn) v = J.in[base + i];
    }
    xs[i] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < spb) {
    const long sys = sys0 + threadIdx.x;
    if (sys < nsys) {
      const int mm = (int)(sys % g.ntchopdim), k = (int)(sys / g.ntchopdim);
      const int nn = mm < g.ntchop ? s_nrchops(g.nrchop, mm * g.minc) : 0;
      if (nn >= 1)
        solve_t_system(g, bxm, at1, op, J.alp, factor_view(J.F, R, ncls), ncls, mm, k, nn, xs + (size_t)threadIdx.x * R,
                       J.ln_in, J.ln_out);
    }
  }
  __syncthreads();
  for (long i = threadIdx.x; i < nv; i += blockDim.x)
    if (sys0 + i / R < nsys) J.out[base + i] = xs[i];
}

// The transposed solve with P lanes per system (band_lanes): A^T = U^T D L^T
// on the partitioned passes, forward over the rows with row n's couplings
// U(n-1, n), U(n-2, n) and scale 1/pivot(n), back with L(n+1, n), L(n+2, n).
// The columns are staged in LDS chopped to NRCHOPS(m) (IDEL2LN's m = 0
// columns with the transpose of the fix at infinity, and at k = 0 the LN
// adjoint in row 0) and written back whole, zeros beyond NRCHOPS(m)
// included.  IHELM's m = 0 systems (their own form, XM^T and LN terms) take
// blocks [0, nm0), one per k, solved by one thread from LDS as in
// k_band_apply_t.  in may equal out.
__global__ void __launch_bounds__(64, 3) k_band_apply_t_part(KGrid g, const double *bxm, const double *at1, int op,
                                                             BandTJobs jobs, int nm0, int npart, int P) {
  extern __shared__ __align__(16) double2 xs[];
  constexpr int LS = PART_LS, SLOT = PART_LS + 1;
  const BandTJob &J = jobs.j[blockIdx.y];
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long ncls = (long)g.ntchopdim * g.nxchop;
  const int R = g.nrchopdim, lane = threadIdx.x;
  if ((int)blockIdx.x < nm0) {
    // IHELM's m = 0 system of slot k on all 64 lanes (lane l owns rows
    // 8 l .. 8 l + 7): XM^T, then (U^T D L^T) on NN-1 rows with three
    // couplings per row, row NN-1 zero, then the LN term's adjoint
    const int k = blockIdx.x;
    const size_t base = (size_t)R * ((size_t)g.ntchopdim * k);
    const int nn = g.ntchop > 0 ? s_nrchops(g.nrchop, 0) : 0;
    auto slot = [&](int n) { return (n / LS) * SLOT + n % LS; };
    if (nn >= 1) {
      for (int n = lane; n < nn; n += 64) xs[slot(n)] = J.in[base + n];
      __syncthreads();
      // x <- XM^T x on NN rows, from the original values (solve_t_system's sums)
      double2 o[LS];
#pragma unroll
      for (int q = 0; q < LS; q++) {
        const int c = lane * LS + q;
        o[q] = zero2();
        if (c < nn) {
          double2 t = axc(bxm[(size_t)c * 3 + 1], xs[slot(c)]);
          if (c + 1 < nn) t = add2(t, axc(bxm[(size_t)(c + 1) * 3 + 0], xs[slot(c + 1)]));
          if (c >= 1) t = add2(t, axc(bxm[(size_t)(c - 1) * 3 + 2], xs[slot(c - 1)]));
          o[q] = t;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < LS; q++)
        if (lane * LS + q < nn) xs[slot(lane * LS + q)] = o[q];
      __syncthreads();
      const int nd = nn - 1;
      const double *f7 = factor_view(J.F, R, ncls).f7;
      const long cls = (long)g.ntchopdim * (k < g.nxchop ? k : g.nxchopdim - k);
      const int nlast = max(nd - 1, 0);
      auto rec = [&](int n) { return f7 + ((size_t)min(max(n, 0), nlast) * ncls + cls) * FSTRIDE; };
      part_passes<1, 3>(
          xs + (size_t)lane * SLOT, 0, lane * LS, nd, 64, lane,
          [&](int n, double(&c)[3], double &sc) {
#pragma unroll
            for (int t = 0; t < 3; t++) c[t] = (n >= t + 1 && n < nd) ? rec(n - 1 - t)[3 + t] : 0.0;
            sc = rec(n)[6];
          },
          [&](int n, double(&c)[3]) {
#pragma unroll
            for (int t = 0; t < 3; t++) c[t] = n + t + 1 < nd ? rec(n + 1 + t)[t] : 0.0;
          });
      __syncthreads();
      if (lane == 0 && k == 0) {
        const double2 c0 = xs[slot(0)], c1 = nn > 1 ? xs[slot(1)] : zero2(), c2 = nn > 2 ? xs[slot(2)] : zero2();
        const double lbar = J.ln_in[0] - (c0.x * ((4.0 / 3 / g.ell2) / g.norm0) + c1.x * (-2 / g.norm1 / g.ell2) +
                                          c2.x * (+2.0 / 3 / g.norm2 / g.ell2));
        J.ln_out[0] = -lbar / J.alp;
      }
    }
    for (int n = lane; n < R; n += 64) J.out[base + n] = n < nn - 1 ? xs[slot(n)] : zero2();
    return;
  }
  const int W = 64 / P;
  // XCD-aware from the next multiple of 8 on (as k_band_apply_part)
  const int pb = xcd_logical(npart, xcd_grid(nm0));
  if (pb < 0) return;
  const long sys0 = (long)pb * W;
  const size_t fstride = (size_t)64 * SLOT;
  for (int jp = 0; jp < W; jp++) {
    int m2, k2;
    const int ndp = penta_rows(g, op, sys0 + jp, m2, k2);
    const size_t colp = (size_t)R * ((size_t)m2 + (size_t)g.ntchopdim * k2);
    const bool fix = op == 1 && m2 == 0 && ndp > 0;
    const double2 a0 = fix ? J.in[colp] : zero2();
    for (int n = lane; n < ndp; n += 64) {
      double2 v = J.in[colp + n];
      if (fix) {
        v = sub2(v, axc(at1[n] / at1[0], a0));
        if (k2 == 0 && n == 0) v = make_double2(g.norm0 * J.ln_in[0], 0.0);
      }
      xs[(jp * P + n / LS) * SLOT + n % LS] = v;
    }
  }
  __syncthreads();
  {
    const int j = lane / P, seg = lane % P;
    int mm, k;
    const int nd = penta_rows(g, op, sys0 + j, mm, k);
    const int kabs = k < g.nxchop ? k : g.nxchopdim - k;
    const long cls = mm + (long)g.ntchopdim * kabs;
    const FactorView FV = factor_view(J.F, R, ncls);
    const int nlast = max(nd - 1, 0);
    auto row = [&](int n) { return (size_t)min(max(n, 0), nlast) * ncls + cls; };
    part_passes<1, 2>(
        xs + (size_t)lane * SLOT, fstride, seg * LS, nd, P, seg,
        [&](int n, double(&c)[2], double &sc) {
          c[0] = (n >= 1 && n < nd) ? FV.fb[row(n - 1) * 2] : 0.0;
          c[1] = (n >= 2 && n < nd) ? FV.fb[row(n - 2) * 2 + 1] : 0.0;
          sc = FV.ff[row(n) * 3 + 2];
        },
        [&](int n, double(&c)[2]) {
          c[0] = n + 1 < nd ? FV.ff[row(n + 1) * 3] : 0.0;
          c[1] = n + 2 < nd ? FV.ff[row(n + 2) * 3 + 1] : 0.0;
        });
  }
  __syncthreads();
  for (int jp = 0; jp < W; jp++) {
    const long sys = sys0 + jp;
    if (sys >= nsys) break;
    const int m2 = (int)(sys % g.ntchopdim), k2 = (int)(sys / g.ntchopdim);
    if (solve_kind(op, m2, k2) != 0) continue;  // IHELM's m = 0: its own block
    const int ndp = m2 < g.ntchop ? s_nrchops(g.nrchop, m2 * g.minc) : 0;
    const size_t colp = (size_t)R * (size_t)sys;
    for (int n = lane; n < R; n += 64) J.out[colp + n] = n < ndp ? xs[(jp * P + n / LS) * SLOT + n % LS] : zero2();
  }
}

int launch_band_solve_t_n(mlg_ctx *c, int op, int n, const BandTJob *in) {
  if ((op != 1 && op != 2) || n < 1 || n > MAX_BAND_JOBS) return MLG_E_ARG;
  KGrid g = kgrid(c);
  if (op == 2 && g.nrchop < 4) return MLG_E_ARG;  // the m = 0 form needs the three LN rows
  BandTJobs jobs;
  const double *used[MAX_BAND_JOBS];
  for (int i = 0; i < n; i++) {
    int rc = 0;
    const double *F = factor_set(c, op, in[i].alp, used, i, &rc);
    if (!F) return rc;
    used[i] = F;
    jobs.j[i] = in[i];
    jobs.j[i].F = F;
  }
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  KScope ks(c, K_BAND_SOLVE, n * 2.0 * 16.0 * fff_count(g));
  const int lanes = band_lanes(g);
  if (lanes) {
    const int nm0 = op == 2 ? g.nxchopdim : 0;
    const long npart = (nsys + 64 / lanes - 1) / (64 / lanes);
    const size_t lds = std::max(part_lds_bytes(1), sizeof(double2) * (size_t)g.nrchopdim);
    hipLaunchKernelGGL(k_band_apply_t_part, dim3((unsigned)(xcd_grid(nm0) + xcd_grid((int)npart)), n), dim3(64), lds,
                       c->stream, g, c->d.bxm, c->d.at1, op, jobs, nm0, (int)npart, lanes);
    MLG_CHECK_HIP(hipGetLastError());
    return 0;
  }
  int spb = (int)((48 * 1024) / ((size_t)g.nrchopdim * sizeof(double2)));
  spb = spb < 1 ? 1 : (spb > 64 ? 64 : spb);
  hipLaunchKernelGGL(k_band_apply_t, dim3((unsigned)((nsys + spb - 1) / spb), n), dim3(64),
                     sizeof(double2) * (size_t)spb * g.nrchopdim, c->stream, g, c->d.bxm, c->d.at1, op, jobs, spb);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- IHELMP (VISC%SW = 2)
// A = (del^P + nu del^2 - alp)^-1 B, even P in 4..14 (src/mod_legops.f90:
// 1409-1592).  Per m the band of del_h^2 over NP = NRCHOPS(m) + P rows
// (host-built, BAND_LOGLEG_RAT_DEL2H) is raised to the powers del_h^(2j),
// j <= P/2 (one block per m, one thread per row, BANMUL's band x band order);
// per (m, k) the (2P+1)-band matrix (del_h^2 - k^2)^(P/2) + nu (del_h^2 -
// k^2) - alp is assembled from the powers with the binomial terms in the
// reference's order and factorised without pivoting (LUDCMPB) over the
// NRCHOPS(m) rows, one thread per system.  The factors are cached like the
// other band solves'.  This path is off the benchmark workloads (VISC%SW = 2
// runs AB2 only), so its kernels are plain one-thread-per-system loops.
constexpr int HELMP_WMAX = 29;  // 2P+1 for P = 14

__global__ void __launch_bounds__(256) k_helmp_powers(int q, int npmax, int nrchop, int minc, const double *__restrict__ d1,
                                                      double *__restrict__ dp) {
  const int mm = blockIdx.x, w = 4 * q + 1;
  const int n = s_nrchops(nrchop, mm * minc) + 2 * q;
  const double *b = d1 + (size_t)mm * npmax * 5;
  double *lev = dp + (size_t)mm * (q + 1) * npmax * w;  // level j: del_h^(2j), width 4j+1, centre 2j
  for (int r = threadIdx.x; r < npmax; r += blockDim.x)
    for (int c = 0; c < w; c++) lev[((size_t)npmax + r) * w + c] = (c < 5 && r < n) ? b[(size_t)r * 5 + c] : 0.0;
  __syncthreads();
  for (int j = 2; j <= q; j++) {
    const double *a = lev + (size_t)(j - 1) * npmax * w;
    double *out = lev + (size_t)j * npmax * w;
    const int aw = 4 * j - 3, ac = 2 * j - 2;
    for (int r = threadIdx.x; r < npmax; r += blockDim.x) {
      double acc[HELMP_WMAX];
      for (int c = 0; c < w; c++) acc[c] = 0.0;
      if (r < n)
        for (int i = 0; i < aw; i++) {
          const int as = i - ac;
          if (r < max(0, -as) || r > min(n - 1, n - 1 - as)) continue;
          for (int jj = 0; jj < 5; jj++) {
            const int bs = jj - 2, rb = r + as;
            if (rb < max(0, -bs) || rb > min(n - 1, n - 1 - bs)) continue;
            acc[i + jj] += a[(size_t)r * w + i] * b[(size_t)rb * 5 + jj];
          }
        }
      for (int c = 0; c < w; c++) out[(size_t)r * w + c] = acc[c];
    }
    __syncthreads();
  }
}

__constant__ double c_binom[8][8] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1}, {1, 5, 10, 10, 5, 1},
                                    {1, 6, 15, 20, 15, 6, 1}, {1, 7, 21, 35, 35, 21, 7, 1}};

// One thread per (m, k): the system's factors F[((sys NRCHOP) + row) W + col]
__global__ void __launch_bounds__(64) k_helmp_factor(KGrid g, int p, int npmax, const double *__restrict__ dp,
                                                     const double *__restrict__ ak, double alp, double nu,
                                                     double *__restrict__ F) {
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long sys = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const int mm = (int)(sys % g.ntchopdim), kk = (int)(sys / g.ntchopdim);
  const int nn = s_nrchops(g.nrchop, mm * g.minc);
  if (nn < 1 || mm >= g.ntchop) return;
  const int q = p / 2, w = 2 * p + 1;
  const double *lev = dp + (size_t)mm * (q + 1) * npmax * w;
  const double akv = ak[mm + (size_t)g.ntchopdim * kk], k2 = akv * akv;
  double *h = F + (size_t)sys * g.nrchop * w;
  for (int r = 0; r < nn; r++) {
    double *hr = h + (size_t)r * w;
    const double *dq = lev + ((size_t)q * npmax + r) * w;
    for (int c = 0; c < w; c++) hr[c] = dq[c];
    for (int t = 1; t < q; t++) {
      const int j = q - t, wj = 4 * j + 1, off = 2 * (q - j);
      double kp = k2;
      for (int e = 1; e < t; e++) kp = kp * k2;
      const double cf = (t & 1) ? -(c_binom[q][t] * kp) : c_binom[q][t] * kp;
      const double *dj = lev + ((size_t)j * npmax + r) * w;
      for (int c = 0; c < wj; c++) hr[off + c] = hr[off + c] + cf * dj[c];
    }
    double kp = k2;
    for (int e = 1; e < q; e++) kp = kp * k2;
    hr[p] = (q & 1) ? hr[p] - kp : hr[p] + kp;
    const double *d1 = lev + ((size_t)npmax + r) * w;
    for (int c = 0; c < 5; c++) {
      double dv = d1[c];
      if (c == 2) dv = dv - k2;
      hr[p - 2 + c] = hr[p - 2 + c] + nu * dv;
    }
    hr[p] = hr[p] - alp;
  }
  // LUDCMPB (src/mod_bandmat.f90:65-156) over nn rows, centre column p
  const int c0 = p;
  for (int r = 0; r < nn; r++) {
    double *hr = h + (size_t)r * w;
    for (int b = c0 + 1; b < w; b++)
      if (r + (b - c0) < nn) hr[b] /= hr[c0];
    for (int b = c0 + 1; b < w; b++)
      for (int j = 1; j <= c0; j++)
        if (r + max(b - c0, j) < nn) h[(size_t)(r + j) * w + (b - j)] -= hr[b] * h[(size_t)(r + j) * w + (c0 - j)];
    for (int j = 1; j <= c0; j++)
      if (r + j < nn) h[(size_t)(r + j) * w + (c0 - j)] /= hr[c0];
  }
  for (int r = 0; r < nn; r++) h[(size_t)r * w + c0] = 1.0 / h[(size_t)r * w + c0];
}

// LUBKSBB (src/mod_bandmat.f90:320-402) on each (m, k) column of A, in place
__global__ void __launch_bounds__(64) k_helmp_solve(KGrid g, int p, const double *__restrict__ F, double2 *__restrict__ a) {
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const long sys = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (sys >= nsys) return;
  const int mm = (int)(sys % g.ntchopdim), kk = (int)(sys / g.ntchopdim);
  const int nn = s_nrchops(g.nrchop, mm * g.minc);
  if (nn < 1 || mm >= g.ntchop) return;
  const int w = 2 * p + 1, c0 = p;
  const double *h = F + (size_t)sys * g.nrchop * w;
  double2 *x = a + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * kk);
  for (int r = 1; r < nn; r++) {
    const int ne = min(r, c0);
    double2 v = x[r];
    for (int j = 1; j <= ne; j++) v = sub2(v, axc(h[(size_t)r * w + (c0 - j)], x[r - j]));
    x[r] = v;
  }
  for (int r = 0; r < nn; r++) x[r] = axc(h[(size_t)r * w + c0], x[r]);
  for (int r = nn - 2; r >= 0; r--) {
    double2 v = x[r];
    for (int i = 1; i <= w - 1 - c0; i++)
      if (r + i < nn) v = sub2(v, axc(h[(size_t)r * w + (c0 + i)], x[r + i]));
    x[r] = v;
  }
}

// A%LN = -B%LN / alp and the log term's contribution at m = 0, k = 0:
// AL2 = LN (4/(3 L^2) / NORM(0), -2 / NORM(1) / L^2, 2/3 / NORM(2) / L^2),
// AL = (del_h^2)^(P/2-1) AL2 (BANMUL), A(:,0,0) -= AL + nu AL2.  One block.
__global__ void __launch_bounds__(256) k_helmp_logterm(KGrid g, int p, int npmax, const double *__restrict__ d1,
                                                       const double *__restrict__ ln_b, double *__restrict__ ln_a,
                                                       double alp, double nu, double2 *__restrict__ a) {
  extern __shared__ double hl[];  // al2[nrchop], al[nrchop], t[nrchop]
  const int nr = g.nrchop;
  double *al2 = hl, *al = hl + nr, *t = hl + 2 * nr;
  const double ln = -ln_b[0] / alp;
  for (int n = threadIdx.x; n < nr; n += blockDim.x) al2[n] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    ln_a[0] = ln;
    if (nr > 0) al2[0] = ln * (4.0 / 3 / g.ell2) / g.norm0;
    if (nr > 1) al2[1] = ln * (-2 / g.norm1 / g.ell2);
    if (nr > 2) al2[2] = ln * (+2.0 / 3 / g.norm2 / g.ell2);
  }
  __syncthreads();
  for (int n = threadIdx.x; n < nr; n += blockDim.x) al[n] = al2[n];
  __syncthreads();
  for (int it = 4; it <= p; it += 2) {
    for (int cc = threadIdx.x; cc < nr; cc += blockDim.x) {
      const double *br = d1 + (size_t)cc * 5;  // m = 0 rows
      double v = br[2] * al[cc];
      for (int bi = 0; bi < 5; bi++) {
        if (bi == 2) continue;
        const int is = bi - 2;
        if (cc + is >= 0 && cc + is <= nr - 1) v += br[bi] * al[cc + is];
      }
      t[cc] = v;
    }
    __syncthreads();
    for (int n = threadIdx.x; n < nr; n += blockDim.x) al[n] = t[n];
    __syncthreads();
  }
  for (int n = threadIdx.x; n < nr; n += blockDim.x) a[n].x = (a[n].x - al[n]) - nu * al2[n];
  (void)npmax;  // the m = 0 band rows are the first of the table
}

int launch_ihelmp(mlg_ctx *c, int p, const double2 *b, const double *ln_b, double2 *a, double *ln_a, double alp, double nu) {
  KGrid g = kgrid(c);
  if (p < 4 || p > 14 || (p & 1) || alp == 0.0) return 3;
  // the log-term block stages 3 NRCHOP doubles in LDS (at most 64 KB)
  if (g.nrchop < 1 || g.nrchop + p > c->g.nnorm || 3L * g.nrchop * 8 > 65536) return MLG_E_ARG;
  const int q = p / 2, w = 2 * p + 1, npmax = g.nrchop + p;
  const long nsys = (long)g.ntchopdim * g.nxchopdim;
  const size_t nfac = (size_t)nsys * g.nrchop * w;  // the set: factors, then the del_h^2 band rows [m][NP][5]
  const double *F = nullptr;
  for (auto &f : c->fcache)
    if (f.op == 100 + p && f.alp == alp && f.nu == nu && f.nrchop == g.nrchop) F = f.f;
  if (!F) {
    std::vector<double> d1((size_t)g.ntchopdim * npmax * 5, 0.0);
    for (int mm = 0; mm < g.ntchopdim; mm++) {
      const int np = nrchops_of(g.nrchop, mm * g.minc) + p;
      band_del2h(np, c->g.m[mm], c->g.p.ell, &c->g.lognorm[(size_t)c->g.nnorm * mm], &d1[(size_t)mm * npmax * 5]);
    }
    double *dp = nullptr, *Fn = nullptr;
    if (hipMalloc((void **)&Fn, sizeof(double) * (nfac + d1.size())) != hipSuccess) return MLG_E_NOMEM;
    if (hipMalloc((void **)&dp, sizeof(double) * (size_t)g.ntchopdim * (q + 1) * npmax * w) != hipSuccess) {
      (void)hipFree(Fn);
      return MLG_E_NOMEM;
    }
    int rc = 0;
    if (hipMemcpyAsync(Fn + nfac, d1.data(), sizeof(double) * d1.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      rc = MLG_E_HIP;
    if (!rc) {
      hipLaunchKernelGGL(k_helmp_powers, dim3(g.ntchopdim), dim3(256), 0, c->stream, q, npmax, g.nrchop, g.minc, Fn + nfac,
                         dp);
      hipLaunchKernelGGL(k_helmp_factor, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, c->stream, g, p, npmax, dp,
                         c->d.ak, alp, nu, Fn);
      if (hipGetLastError() != hipSuccess) rc = MLG_E_HIP;
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dp);
    if (rc) {
      (void)hipFree(Fn);
      return rc;
    }
    FactorSet fs{100 + p, alp, g.nrchop, Fn};
    fs.nu = nu;
    c->fcache.push_back(fs);
    F = Fn;
  }
  if (a != b) {
    const int rc = launch_copy(c, a, b, fff_count(g));
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_helmp_logterm, dim3(1), dim3(256), sizeof(double) * 3 * (size_t)g.nrchop, c->stream, g, p, npmax,
                     F + nfac, ln_b, ln_a, alp, nu, a);
  hipLaunchKernelGGL(k_helmp_solve, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, c->stream, g, p, F, a);
  MLG_CHECK_HIP(hipGetLastError());
  const int rc = launch_fix_at_infinity(c, a);
  return rc ? rc : launch_chopdo(c, a, MLG_FFF);
}

// ---------------------------------------------------------------- pointwise updates
// Pointwise FFF updates with the reference's evaluation orders:
//   mode 0: y = x1 + a*x2                      (Richardson / Euler sub-steps)
//   mode 1: y = x1 + a*(1.5*x2 - 0.5*x3)       (AB2 update)
//   mode 2: y = 2*x1 - y                       (Richardson extrapolation)
//   mode 3: y = a*(x1 - x2)                    (pressure history)
//   mode 4: y = x1 - x2
//   mode 5: y = -y - a*x1                      (Crank-Nicolson right-hand side)
__global__ void k_axpby(int mode, double2 *y, double a, const double2 *x1, const double2 *x2, const double2 *x3, long n) {
  GRID_STRIDE(n) {
    double2 v;
    if (mode == 0) v = add2(x1[idx], axc(a, x2[idx]));
    else if (mode == 1) v = add2(x1[idx], axc(a, sub2(axc(1.5, x2[idx]), axc(0.5, x3[idx]))));
    else if (mode == 2) v = sub2(axc(2.0, x1[idx]), y[idx]);
    else if (mode == 3) v = axc(a, sub2(x1[idx], x2[idx]));
    else if (mode == 4) v = sub2(x1[idx], x2[idx]);
    else v = sub2(make_double2(-y[idx].x, -y[idx].y), axc(a, x1[idx]));
    y[idx] = v;
  }
}
__global__ void k_ln_axpby(int mode, double *y, double a, const double *x1, const double *x2, const double *x3) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v;
  if (mode == 0) v = x1[0] + a * x2[0];
  else if (mode == 1) v = x1[0] + a * (1.5 * x2[0] - 0.5 * x3[0]);
  else if (mode == 2) v = 2.0 * x1[0] - y[0];
  else if (mode == 3) v = a * (x1[0] - x2[0]);
  else v = x1[0] - x2[0];
  y[0] = v;
}
// The middle of STEP_BOUSSI_AB_CN (src/mod_march.f90:609-661) for up to
// three fields in one pass over the FFF elements: Y0 = COPY0(Y) (chopped),
// Y = Y + DT (1.5 X - 0.5 H) (AB2, every element), H = COPY0(X); LN
// likewise (the AB2 update of LN only where ln is set).  Element by element
// these are k_copy0_fff's and k_axpby's operations.
__global__ void k_ab2_hist(KGrid g, Ab2Jobs jobs, double dt, long nfff) {
  const Ab2Job &J = jobs.j[blockIdx.y];
  GRID_STRIDE(nfff) {
    const int n = (int)(idx % g.nrchopdim), mm = (int)((idx / g.nrchopdim) % g.ntchopdim);
    const int nn = mm < g.ntchop ? s_nrchops(g.nrchop, mm * g.minc) : 0;
    const bool in = n < nn;
    const double2 y = J.y[idx], x = J.x[idx], h = J.h[idx];
    J.y0[idx] = in ? y : zero2();
    J.y[idx] = add2(y, axc(dt, sub2(axc(1.5, x), axc(0.5, h))));
    J.h[idx] = in ? x : zero2();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double yl = J.y_ln[0], xl = J.x_ln[0], hl = J.h_ln[0];
    J.y0_ln[0] = yl * 1.0 + 0.0;
    if (J.ln) J.y_ln[0] = yl + dt * (1.5 * xl - 0.5 * hl);
    J.h_ln[0] = xl * 1.0 + 0.0;
  }
}
int launch_ab2_hist(mlg_ctx *c, int n, const Ab2Jobs &jobs, double dt) {
  if (n < 1 || n > 3) return MLG_E_ARG;
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_ab2_hist, dim3(nblocks(c->n_fff), n), dim3(256), 0, c->stream, g, jobs, dt, c->n_fff);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_axpby(mlg_ctx *c, int mode, double2 *y, double a, const double2 *x1, const double2 *x2, const double2 *x3, long n) {
  hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(256), 0, c->stream, mode, y, a, x1, x2, x3, n);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}
int launch_ln_axpby(mlg_ctx *c, int mode, double *y, double a, const double *x1, const double *x2, const double *x3) {
  hipLaunchKernelGGL(k_ln_axpby, dim3(1), dim3(64), 0, c->stream, mode, y, a, x1, x2, x3);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// ETD update (CALC_BOUSSI_ETDFE_SUB / _ETDAB_SUB, src/mod_march.f90:948-1139)
// fused with what surrounds it in STEP_BOUSSI_ETDFE_BE / _ETDAB_CN
// (:664-901): the HYPERV3 filter of the nonlinear terms before the update
// and of the updated fields after it (CALC_BOUSSI_ETD_HYPERV :1352-1416),
// the copy of b before the update, and the stores of the filtered
// nonlinear terms and updated velocities into the solver's history.  Every
// value goes through the same operations as the separate kernels.
struct Mat4x2 {
  double e[16], nl[16];
};
struct EtdArgs {
  double2 *rur, *rup, *uz, *b;       // state: in, updated out (b: in only)
  const double2 *nr, *np, *nz, *nb;  // nonlinear terms (before filtering)
  double2 *hpr, *hpp, *hpz;          // pressure-gradient history: read (ab), written (save & 2)
  double2 *hnr, *hnp, *hnz, *hnb;    // nonlinear-term history: read (ab), written (save & 1)
  double2 *b0;                       // receives b before the update (or null)
  double2 *bo;                       // receives the updated b (may be b)
  double *lnd[3];                    // LN values set on the way: *lnd[i] = *lns[i], or 0 for a null
  const double *lns[3];              // lns[i] (null lnd[i]: none)
  const double *hv, *hd;             // filter tables [hr | ht | hx] (or null: no filter)
  int nrc;
  double tf;
  int post;                          // the filter also acts on the updated u, b
};

// HYPERV3 factor h^tf of one coefficient (0 outside the truncation); tf is
// 1 on full steps, where h^1 = h
__device__ __forceinline__ double filt_factor(const KGrid &g, const FIdx &f, int nn, const double *h, int nrc, double tf) {
  if (f.n < nn && f.mm < g.ntchop) {
    const double *hr = h, *ht = h + (size_t)nrc * g.ntchopdim, *hx = ht + g.ntchopdim;
    double hh = hr[f.n + (size_t)nrc * f.mm] * ht[f.mm] * hx[f.k];
    return tf == 1.0 ? hh : pow(hh, tf);
  }
  return 0.0;
}
__device__ __forceinline__ double2 filt_apply(const KGrid &g, const FIdx &f, int nn, double fac, double2 v) {
  return (f.n < nn && f.mm < g.ntchop) ? axc(fac, v) : zero2();
}

__global__ void __launch_bounds__(256) k_etd_sub(KGrid g, Mat4x2 M, double dt, EtdArgs A, int ab, int save) {
  if (blockIdx.x == 0 && threadIdx.x < 3 && A.lnd[threadIdx.x])
    A.lnd[threadIdx.x][0] = A.lns[threadIdx.x] ? A.lns[threadIdx.x][0] : 0.0;
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    const int nn = s_nrchops(g.nrchop, f.mm * g.minc);
    if (A.hv && A.post && !A.b0 && !(f.n < nn && f.mm < g.ntchop)) {
      // outside the truncation the filter's factor is zero: every value
      // written below is the filtered one (zero), so nothing is read
      A.rur[idx] = zero2();
      A.rup[idx] = zero2();
      A.uz[idx] = zero2();
      A.bo[idx] = zero2();
      if (save & 1) A.hnr[idx] = A.hnp[idx] = A.hnz[idx] = A.hnb[idx] = zero2();
      if (save & 2) A.hpr[idx] = A.hpp[idx] = A.hpz[idx] = zero2();
      continue;
    }
    double er[16], ar[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      er[q] = M.e[q];
      ar[q] = M.nl[q];
    }
    double notk0 = 1.0, notm0 = (f.mm == 0) ? 0.0 : 1.0;
    if (f.k == 0) {
      notk0 = 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        er[0 + 4 * j] = 0.0;
        er[1 + 4 * j] = 0.0;
        ar[0 + 4 * j] = 0.0;
        ar[1 + 4 * j] = 0.0;
      }
      er[0] = 1.0;
      er[5] = 1.0;
      ar[0] = dt;
      ar[5] = dt;
    }
    double2 nl[4] = {A.nr[idx], A.np[idx], A.nz[idx], A.nb[idx]};
    double fv = 0.0, fd = 0.0;  // filter factors (velocity, buoyancy)
    if (A.hv) {
      fv = filt_factor(g, f, nn, A.hv, A.nrc, A.tf);
      fd = filt_factor(g, f, nn, A.hd, A.nrc, A.tf);
#pragma unroll
      for (int q = 0; q < 3; q++) nl[q] = filt_apply(g, f, nn, fv, nl[q]);
      nl[3] = filt_apply(g, f, nn, fd, nl[3]);
    }
    const double2 bold = A.b[idx];
    if (A.b0) A.b0[idx] = bold;
    double2 ev[4], av[4];
    if (ab) {
      ev[0] = sub2(A.rur[idx], A.hpr[idx]);
      ev[1] = sub2(A.rup[idx], axc(notm0, A.hpp[idx]));
      ev[2] = sub2(A.uz[idx], axc(notk0, A.hpz[idx]));
      ev[3] = bold;
      av[0] = sub2(axc(1.5, nl[0]), axc(0.5, A.hnr[idx]));
      av[1] = sub2(axc(1.5, nl[1]), axc(0.5, A.hnp[idx]));
      av[2] = sub2(axc(1.5, nl[2]), axc(0.5, A.hnz[idx]));
      av[3] = sub2(axc(1.5, nl[3]), axc(0.5, A.hnb[idx]));
    } else {
      ev[0] = A.rur[idx];
      ev[1] = A.rup[idx];
      ev[2] = A.uz[idx];
      ev[3] = bold;
      av[0] = nl[0];
      av[1] = nl[1];
      av[2] = nl[2];
      av[3] = nl[3];
    }
    double2 cv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double2 s1 = zero2(), s2 = zero2();
#pragma unroll
      for (int j = 0; j < 4; j++) {
        s1 = add2(s1, axc(er[i + 4 * j], ev[j]));
        s2 = add2(s2, axc(ar[i + 4 * j], av[j]));
      }
      cv[i] = add2(s1, s2);
    }
    if (A.hv && A.post) {
#pragma unroll
      for (int q = 0; q < 3; q++) cv[q] = filt_apply(g, f, nn, fv, cv[q]);
      cv[3] = filt_apply(g, f, nn, fd, cv[3]);
    }
    A.rur[idx] = cv[0];
    A.rup[idx] = cv[1];
    A.uz[idx] = cv[2];
    A.bo[idx] = cv[3];
    if (save & 1) {
      A.hnr[idx] = nl[0];
      A.hnp[idx] = nl[1];
      A.hnz[idx] = nl[2];
      A.hnb[idx] = nl[3];
    }
    if (save & 2) {
      A.hpr[idx] = cv[0];
      A.hpp[idx] = cv[1];
      A.hpz[idx] = cv[2];
    }
  }
}
int launch_etd_sub(mlg_ctx *c, const double *e16, const double *nl16, double dt, double2 *rur, double2 *rup, double2 *uz,
                   double2 *b, const double2 *nr_, const double2 *np_, const double2 *nz_, const double2 *nb_, double2 *hpr,
                   double2 *hpp, double2 *hpz, double2 *hnr, double2 *hnp, double2 *hnz, double2 *hnb, int ab, int save,
                   double2 *b0, double2 *bout, int filter, double tf, double *const *lnd, const double *const *lns,
                   int post_filter) {
  KGrid g = kgrid(c);
  Mat4x2 M;
  for (int q = 0; q < 16; q++) {
    M.e[q] = e16[q];
    M.nl[q] = nl16[q];
  }
  EtdArgs A{rur, rup, uz, b, nr_, np_, nz_, nb_, hpr, hpp, hpz, hnr, hnp, hnz, hnb, b0, bout ? bout : b,
            {lnd ? lnd[0] : nullptr, lnd ? lnd[1] : nullptr, lnd ? lnd[2] : nullptr},
            {lns ? lns[0] : nullptr, lns ? lns[1] : nullptr, lns ? lns[2] : nullptr},
            nullptr, nullptr, 0, tf, post_filter};
  if (filter) {
    const mlg::Grid &G = c->g;
    size_t nr1 = (size_t)G.p.nrchop * G.ntchopdim, nt = G.ntchopdim, nx = G.nxchopdim;
    A.hv = c->d.hv;
    A.hd = c->d.hv + (nr1 + nt + nx);
    A.nrc = G.p.nrchop;
  }
  // 4 fields + 4 nonlinear terms in, 4 out; history: 7 in (ab), 4 + 3 out
  // (save bits); b0 out.  With the filter (and no b0) only the coefficients
  // inside the truncation are read.
  const double nin = 8.0 + (ab ? 7.0 : 0.0), nout = 4.0 + ((save & 1) ? 4.0 : 0.0) + ((save & 2) ? 3.0 : 0.0) + (b0 ? 1.0 : 0.0);
  const double cin = (A.hv && A.post && !b0) ? (double)sum_nrchops(c) * g.nxchopdim : (double)fff_count(g);
  KScope ks(c, K_ETD, 16.0 * (cin * nin + (double)fff_count(g) * nout));
  hipLaunchKernelGGL(k_etd_sub, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, M, dt, A, ab, save);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// HYPERV3 / CALC_BOUSSI_ETD_HYPERV filter (src/mod_march.f90:1352-1416,3521-3556)
__global__ void k_hyp3(KGrid g, int nrc, const double *hr, const double *ht, const double *hx, double tf, double2 *a) {
  GRID_STRIDE(fff_count(g)) {
    FIdx f = fdecode(g, idx);
    int nn = s_nrchops(g.nrchop, f.mm * g.minc);
    if (f.n < nn && f.mm < g.ntchop) {
      double h = hr[f.n + (size_t)nrc * f.mm] * ht[f.mm] * hx[f.k];
      a[idx] = axc(pow(h, tf), a[idx]);
    } else {
      a[idx] = zero2();
    }
  }
}
int launch_hyp3(mlg_ctx *c, double2 *a, int diff, double tf) {
  KGrid g = kgrid(c);
  const mlg::Grid &G = c->g;
  int nrc = G.p.nrchop;
  size_t nr_ = (size_t)nrc * G.ntchopdim, nt = G.ntchopdim, nx = G.nxchopdim;
  const double *base = c->d.hv + (diff ? (nr_ + nt + nx) : 0);
  hipLaunchKernelGGL(k_hyp3, dim3(nblocks(fff_count(g))), dim3(256), 0, c->stream, g, nrc, base, base + nr_, base + nr_ + nt, tf,
                     a);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// DIAGNOST per-step checks (src/mod_march.f90:1750-1791): dealiasing of psi
// beyond NRCHOPS at m = 0, 1 (error 11) and |psi| > 100 in the low modes
// n < 5, 1 <= m < 6, k < 5 (error 12).
__global__ void __launch_bounds__(256) k_diagnost(KGrid g, const double2 *psi, int *err) {
  __shared__ double m0[256], m1[256];
  __shared__ int div;
  if (threadIdx.x == 0) div = 0;
  double a0 = -1e300, a1 = -1e300;
  const int n0 = s_nrchops(g.nrchop, 0), n1 = s_nrchops(g.nrchop, g.minc);
  const int r0 = g.nrchopdim - n0, r1 = g.nrchopdim - n1;  // rows checked per k at m = 0, 1
  const size_t colk = (size_t)g.nrchopdim * (size_t)g.ntchopdim;
  for (int i = threadIdx.x; i < r0 * g.nxchopdim; i += blockDim.x) {
    const int k = i / r0, n = n0 + (i - k * r0);
    double2 v = psi[(size_t)n + colk * k];
    a0 = fmax(a0, hypot(v.x, v.y));
  }
  if (g.ntchopdim >= 2)
    for (int i = threadIdx.x; i < r1 * g.nxchopdim; i += blockDim.x) {
      const int k = i / r1, n = n1 + (i - k * r1);
      double2 v = psi[(size_t)n + (size_t)g.nrchopdim + colk * k];
      a1 = fmax(a1, hypot(v.x, v.y));
    }
  m0[threadIdx.x] = a0;
  m1[threadIdx.x] = a1;
  __syncthreads();
  int kmax = g.nxchopdim < 5 ? g.nxchopdim : 5, mmax = g.ntchopdim < 6 ? g.ntchopdim : 6, nmax = g.nrchop < 5 ? g.nrchop : 5;
  for (int i = threadIdx.x; i < kmax * mmax * nmax; i += blockDim.x) {
    int n = i % nmax, mm = (i / nmax) % mmax, k = i / (nmax * mmax);
    if (mm < 1) continue;
    double2 v = psi[(size_t)n + (size_t)g.nrchopdim * ((size_t)mm + (size_t)g.ntchopdim * k)];
    if (hypot(v.x, v.y) > 100.) div = 1;
  }
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      m0[threadIdx.x] = fmax(m0[threadIdx.x], m0[threadIdx.x + s]);
      m1[threadIdx.x] = fmax(m1[threadIdx.x], m1[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double chk = (g.ntchopdim >= 2) ? (m0[0] + m1[0]) : m0[0];
    if (chk > 1.0e-12) atomicCAS(err, 0, 11);
    else if (div) atomicCAS(err, 0, 12);
  }
}
int launch_diagnost(mlg_ctx *c, const double2 *psi) {
  KGrid g = kgrid(c);
  hipLaunchKernelGGL(k_diagnost, dim3(1), dim3(256), 0, c->stream, g, psi, c->d.err);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// PRODCT_MK_ALLK / L2NORM_MK_ALLK (src/mod_scalar3.f90:470-526,750-803):
// per (m, k) Gauss sum over r in PFF, x 4 pi ZLEN L^2, m = 0 halved.
__global__ void __launch_bounds__(64) k_energy_terms(KGrid g, double fac, const double *w, const double *x, const double2 *a,
                                                     const double2 *b, int l2, double *out) {
  const int mm = blockIdx.x, k = blockIdx.y;
  const size_t cb = (size_t)g.ndimr * ((size_t)mm + (size_t)g.ntchopdim * k);
  double s = 0.0;
  for (int r = threadIdx.x; r < g.nr; r += 64) {
    double2 av = a[cb + r];
    double wk;
    if (l2) {
      double ab = hypot(av.x, av.y), om = 1 - x[r];
      wk = ab * ab / (om * om);
    } else {
      double2 bv = b[cb + r];
      wk = av.x * bv.x + av.y * bv.y;
    }
    s += wk * w[r];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) out[mm + (size_t)g.ntchop * k] = (mm == 0 ? 0.5 : 1.0) * (fac * s);
}
int launch_energy_terms(mlg_ctx *c, const double2 *a, const double2 *b, int l2, double *out_all) {
  KGrid g = kgrid(c);
  const double pi = 3.14159265358979323846;
  double fac = 4.0 * pi * c->g.zlen * c->g.ell2;
  hipLaunchKernelGGL(k_energy_terms, dim3(g.ntchopdim, g.nxchopdim), dim3(64), 0, c->stream, g, fac, c->d.w, c->d.x, a, b, l2,
                     out_all);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace mlg
