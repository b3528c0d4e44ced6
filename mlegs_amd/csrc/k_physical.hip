// Physical-space (PPP) pointwise kernels.
//
// VPRODSUB (src/mod_legops.f90:1972-2133) and VDOTPSUB (:2136-2232) act on
// the real and imaginary parts of a PPP slot independently (they are two
// physical theta points).  The fused FORCE variant also applies the linear
// terms of CALC_BOUSSI_FORCE (src/mod_boussinesq.f90:382-391):
//   r N_r += 2 Omega r u_theta, r N_theta -= 2 Omega r u_r, N_z -= b,
//   N_b = -(u . grad b) + Nbar^2 u_z,
// so the nine velocity/vorticity/gradient fields are read once and the four
// right-hand sides written once.  The free-stream/strain additions and the
// CFL monitor of VPRODSUB are zero/off on this path.
#include <mutex>

#include "mlegs_internal.h"

namespace mlg {

__global__ void k_vprod(KGrid g, const double *__restrict__ rr, const double2 *__restrict__ rur,
                        const double2 *__restrict__ rup, const double2 *__restrict__ uz, double2 *__restrict__ ror,
                        double2 *__restrict__ rop, double2 *__restrict__ oz, long total) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const unsigned u = (unsigned)idx, q = u / (unsigned)g.nr, zz = q / (unsigned)g.nth;
    const int r = (int)(u - q * (unsigned)g.nr), j = (int)(q - zz * (unsigned)g.nth), z = (int)zz;
    size_t id = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    double2 A = rur[id], B = rup[id], Cc = uz[id], D = ror[id], E = rop[id], F = oz[id];
    double r2 = rr[r];  // r_i^2
    ror[id] = make_double2(B.x * F.x - Cc.x * E.x, B.y * F.y - Cc.y * E.y);
    rop[id] = make_double2(Cc.x * D.x - A.x * F.x, Cc.y * D.y - A.y * F.y);
    oz[id] = make_double2((A.x * E.x - B.x * D.x) / r2, (A.y * E.y - B.y * D.y) / r2);
  }
}

__global__ void k_vdot(KGrid g, const double *__restrict__ rr, const double2 *__restrict__ rar,
                       const double2 *__restrict__ rap, const double2 *__restrict__ az, const double2 *__restrict__ rbr,
                       const double2 *__restrict__ rbp, const double2 *__restrict__ bz, double2 *__restrict__ out,
                       double sign, long total) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const unsigned u = (unsigned)idx, q = u / (unsigned)g.nr, zz = q / (unsigned)g.nth;
    const int r = (int)(u - q * (unsigned)g.nr), j = (int)(q - zz * (unsigned)g.nth), z = (int)zz;
    size_t id = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    double2 a1 = rar[id], a2 = rap[id], a3 = az[id], b1 = rbr[id], b2 = rbp[id], b3 = bz[id];
    double r2 = rr[r];  // r_i^2
    double vr = (a1.x * b1.x + a2.x * b2.x) / r2 + a3.x * b3.x;
    double vi = (a1.y * b1.y + a2.y * b2.y) / r2 + a3.y * b3.y;
    out[id] = make_double2(sign * vr, sign * vi);
  }
}

__global__ void k_force_products(KGrid g, const double *__restrict__ rr, const double2 *__restrict__ rur,
                                 const double2 *__restrict__ rup, const double2 *__restrict__ uz,
                                 double2 *__restrict__ ror, double2 *__restrict__ rop, double2 *__restrict__ oz,
                                 const double2 *__restrict__ rbr, const double2 *__restrict__ rbp,
                                 const double2 *__restrict__ bz, const double2 *__restrict__ bphys,
                                 double2 *__restrict__ bn, double two_omega, double n2, int linear, long total) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const unsigned u = (unsigned)idx, q = u / (unsigned)g.nr, zz = q / (unsigned)g.nth;
    const int r = (int)(u - q * (unsigned)g.nr), j = (int)(q - zz * (unsigned)g.nth), z = (int)zz;
    size_t id = (size_t)r + (size_t)g.ndimr * ((size_t)j + (size_t)g.ndimth * z);
    double2 A = rur[id], B = rup[id], Cc = uz[id], D = ror[id], E = rop[id], F = oz[id];
    double2 P = rbr[id], Q = rbp[id], R = bz[id];
    double r2 = rr[r];  // r_i^2
    double2 xr = make_double2(B.x * F.x - Cc.x * E.x, B.y * F.y - Cc.y * E.y);
    double2 xp = make_double2(Cc.x * D.x - A.x * F.x, Cc.y * D.y - A.y * F.y);
    double2 xz = make_double2((A.x * E.x - B.x * D.x) / r2, (A.y * E.y - B.y * D.y) / r2);
    double vr = (A.x * P.x + B.x * Q.x) / r2 + Cc.x * R.x;
    double vi = (A.y * P.y + B.y * Q.y) / r2 + Cc.y * R.y;
    double2 nb = make_double2(-vr, -vi);
    if (linear) {
      double2 w = bphys[id];
      xr = make_double2(xr.x + two_omega * B.x, xr.y + two_omega * B.y);
      xp = make_double2(xp.x - two_omega * A.x, xp.y - two_omega * A.y);
      xz = make_double2(xz.x - w.x, xz.y - w.y);
      nb = make_double2(nb.x + n2 * Cc.x, nb.y + n2 * Cc.y);
    }
    ror[id] = xr;
    rop[id] = xp;
    oz[id] = xz;
    bn[id] = nb;
  }
}

static int grid_for(long total) {
  long b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  return (int)(b > 0 ? b : 1);
}

int launch_vprod(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *ror, double2 *rop, double2 *oz) {
  KGrid g = kgrid(c);
  long total = (long)g.nr * g.nth * g.ndimx;
  KScope ks(c, K_PRODUCTS, 16.0 * total * 9);
  hipLaunchKernelGGL(k_vprod, dim3(grid_for(total)), dim3(256), 0, c->stream, g, c->d.r, rur, rup, uz, ror, rop, oz, total);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_vdot(mlg_ctx *c, const double2 *rar, const double2 *rap, const double2 *az, const double2 *rbr, const double2 *rbp,
                const double2 *bz, double2 *out, double sign) {
  KGrid g = kgrid(c);
  long total = (long)g.nr * g.nth * g.ndimx;
  KScope ks(c, K_PRODUCTS, 16.0 * total * 7);
  hipLaunchKernelGGL(k_vdot, dim3(grid_for(total)), dim3(256), 0, c->stream, g, c->d.r, rar, rap, az, rbr, rbp, bz, out, sign,
                     total);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_force_products(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *ror, double2 *rop,
                          double2 *oz, const double2 *rbr, const double2 *rbp, const double2 *bz, const double2 *bphys,
                          double2 *bn, double two_omega, double n2, int linear) {
  KGrid g = kgrid(c);
  long total = (long)g.nr * g.nth * g.ndimx;
  KScope ks(c, K_PRODUCTS, 16.0 * total * (linear ? 14 : 13));  // 9 (+1) fields in, 4 out
  hipLaunchKernelGGL(k_force_products, dim3(grid_for(total)), dim3(256), 0, c->stream, g, c->d.r, rur, rup, uz, ror, rop, oz,
                     rbr, rbp, bz, bphys, bn, two_omega, n2, linear, total);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

// VEL2PLN (src/mod_legops.f90:2277-2339): the far-field r u_theta of the
// (m=0, k=0) column fixes psi's log coefficient, LN = -u(inf)/2 with
// u(inf) = sum_n a_n P_n(1) of the unnormalised expansion; that log image
// is removed from the physical column.
// Pass 1: one wave per n forms the projection a_n P_n(1) (the parity-folded
// Gauss sum of Re(r u_theta) against the table row, lanes over the nodes).
__global__ void __launch_bounds__(256) k_vel2pln_proj(KGrid g, const double *__restrict__ pf, const double *__restrict__ w,
                                                      const double *__restrict__ lognorm, const double2 *__restrict__ rup,
                                                      double *__restrict__ proj) {
  const int n = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nn = g.nrchop;  // NRCHOPS(1) at m = 0
  if (n >= nn) return;
  const double *pfn = pf + (size_t)n * g.nrh;
  double cr = 0.0;
  for (int i = lane; i < g.nrh; i += 64) {
    const double p = rup[i].x, q = rup[g.nr - 1 - i].x;
    cr += ((n & 1) ? (p - q) : (p + q)) * w[i] * pfn[i];
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) cr += __shfl_xor(cr, o, 64);
  if (lane == 0) proj[n] = cr * exp(lognorm[n]);
}
// Pass 2 (one block of 256): LN = -(sum of the projections) / 2, and
// the log image -LN P_n(x) (n = 0, 1) leaves the physical column.
__global__ void __launch_bounds__(256) k_vel2pln_apply(KGrid g, const double *__restrict__ pf, const double *__restrict__ proj,
                                                       double2 *__restrict__ rup, double *__restrict__ psi_ln, int ln_mode,
                                                       double ln) {
  __shared__ double part[4];
  double v = 0.0;
  for (int n = threadIdx.x; n < g.nrchop; n += 256) v += proj[n];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  const double pln = -(((part[0] + part[1]) + part[2]) + part[3]) / 2.0;
  for (int i = threadIdx.x; i < g.nrh; i += blockDim.x) {
    double e = -pf[i] * pln / g.norm0;
    double o = -pf[(size_t)g.nrh + i] * pln / g.norm1;
    double2 a = rup[i], b = rup[g.nr - 1 - i];
    rup[i] = make_double2(a.x - (e + o), a.y);
    rup[g.nr - 1 - i] = make_double2(b.x - (e - o), b.y);
  }
  if (threadIdx.x == 0 && ln_mode != 2) psi_ln[0] = ln_mode == 1 ? ln : pln;
}

static int vel2pln_passes(mlg_ctx *c, double2 *rup, double *psi_ln, int ln_mode, double ln, double *proj) {
  KGrid g = kgrid(c);
  if (g.nrchop > 0)
    hipLaunchKernelGGL(k_vel2pln_proj, dim3((g.nrchop + 3) / 4), dim3(256), 0, c->stream, g, c->d.pf, c->d.w, c->d.lognorm,
                       rup, proj);
  hipLaunchKernelGGL(k_vel2pln_apply, dim3(1), dim3(256), 0, c->stream, g, c->d.pf, proj, rup, psi_ln, ln_mode, ln);
  MLG_CHECK_HIP(hipGetLastError());
  return 0;
}

int launch_vel2pln(mlg_ctx *c, double2 *rup, double *psi_ln, int ln_mode, double ln) {
  if (!rup || !psi_ln) return MLG_E_ARG;
  return vel2pln_passes(c, rup, psi_ln, ln_mode, ln, c->d.vproj);
}

// VEL2PLN without allocation.  Its one scratch buffer, the NRCHOP
// projections of pass 1, is this fixed device array instead of the context's
// buffer.  Every GPU holds its own copy until the process ends; the copy's
// address is looked up once per device, when the first context is created on
// it, and only that address is handed to kernels.  Pass 1 writes every entry
// that pass 2 reads, so nothing needs clearing between calls; the array keeps
// the last call's projections.  Calls that overlap on one GPU (two streams,
// two threads) share the array and race.
static __device__ __align__(256) double scratch[VEL2PLN_SCRATCH_MAX];
static double *scratch_dev[MLG_MAX_DEVICES];
static std::mutex scratch_mu;

int vel2pln_scratch_setup(int device) {
  if (device < 0 || device >= MLG_MAX_DEVICES) return MLG_E_ARG;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return 0;
  void *ptr = nullptr;
  // the caller has made `device` current
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MLG_E_HIP;
  scratch_dev[device] = (double *)ptr;
  return 0;
}

int launch_vel2pln_noalloc(mlg_ctx *c, double2 *rup, double *psi_ln, int ln_mode, double ln) {
  if (!rup || !psi_ln) return MLG_E_ARG;
  if (c->g.nrchop > VEL2PLN_SCRATCH_MAX) return MLG_E_SIZE;
  double *proj = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (c->device >= 0 && c->device < MLG_MAX_DEVICES) proj = scratch_dev[c->device];
  }
  if (!proj) return MLG_E_HIP;
  return vel2pln_passes(c, rup, psi_ln, ln_mode, ln, proj);
}

}  // namespace mlg
