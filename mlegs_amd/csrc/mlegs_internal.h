// Internal structures of libmlegs_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <unordered_set>
#include <vector>

#include "mlegs_hip.h"

namespace mlg {

// Host-side grid and tables (LEGINIT, src/mod_fft.f90:180-367).
struct Grid {
  mlg_params p;
  int nr, nrh, nth, nx, ntchop, nxchop;
  double ell, ell2, zlen;
  int ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim, nxchoph;
  int nrchop;  // current truncation (CHOPSET state)
  int npf, nnorm, nband;
  std::vector<int> m;
  std::vector<double> ak, x, w, r, lnx, norm, lognorm, pf, at0, at1;
  std::vector<double> bxm, bxxdx, bdel2h;  // [mm][row][3|5]
  std::vector<double> vdt;                 // [3][mm][n][j]
  std::vector<double> hv_r, hv_t, hv_x, hd_r, hd_t, hd_x;
};

int build_grid(const mlg_params &p, Grid &g);
// HYP3SET (src/mod_march.f90:3320-3380) with per-direction coefficients
void hyp3set(const Grid &g, double nu_r, double nu_t, double nu_x, std::vector<double> &hr, std::vector<double> &ht,
             std::vector<double> &hx);
void twiddles(int n, int sign, double *re, double *im);
void band_xm(int ni, int m, const double *ln, double *b);
void band_xxdx(int ni, int m, const double *ln, double *b);
void band_del2h(int ni, int m, double ell, const double *ln, double *b);

// Device copies of the tables and scratch.
struct DevTables {
  double *pf = nullptr;       // [mm][n][i]  (i fastest), npf rows
  double *vdt = nullptr;      // [3][mm][n][j]
  double *w = nullptr, *r = nullptr, *x = nullptr, *lnx = nullptr;
  double *ones = nullptr;     // NR ones: RTRAN forward's Gauss sums without weights (the adjoint's RTRAN backward^T)
  double *at1 = nullptr, *ak = nullptr, *norm = nullptr, *lognorm = nullptr;
  double *bxm = nullptr, *bxxdx = nullptr, *bdel2h = nullptr;
  double2 *tw_z = nullptr;    // e^{-2 pi i j / NX}, j < NX
  double2 *tw_t = nullptr;    // e^{-2 pi i j / NTH}, j < NTH
  double2 *tw_h = nullptr;    // e^{-2 pi i j / (2 NTH)}, j <= NTH
  double2 *tw_h2 = nullptr;   // e^{-2 pi i j / (2 NTH)}, j < 2 NTH (theta DFTs of other lengths)
  double *hv = nullptr;       // [hv_r | hv_t | hv_x | hd_r | hd_t | hd_x]
  double *red = nullptr;      // reduction scratch
  double *vproj = nullptr;    // VEL2PLN projections [NRCHOPDIM]
  double *cvdt = nullptr;     // PROJECT tables composed with RTRAN (built on first use)
  // Split hi/lo FP32 forms of the Legendre tables (built when the split
  // precision is first selected): every row of a product's table operand is
  // scaled by 2^-E, E the row's exponent (|entries| < 2^E), and split into
  // hi (a multiple of 2^-10) and lo (the FP32 rounding of the rest).
  float2 *pf_sb = nullptr;    // PF, rows (m, node i): RTRAN backward
  int *pf_eb = nullptr;
  float2 *pf_sf = nullptr;    // PF, rows (m, n): RTRAN forward
  int *pf_ef = nullptr;
  float2 *vdt_s = nullptr;    // V, D, T, rows (table, m, n): PROJECT
  int *vdt_e = nullptr;
  float2 *cvdt_s = nullptr;   // composed V, D, T, rows (table, m, parity, h)
  int *cvdt_e = nullptr;
  int *err = nullptr;         // sticky device error code
};

}  // namespace mlg

struct mlg_field {
  mlg_ctx *ctx;
  double2 *e;   // device elements (sized for the largest space)
  double *ln;   // device log-term coefficient
  int space;
  bool pooled;
};

namespace mlg {
// Kernel classes timed by the optional event timer.
enum KernelId {
  K_RTRAN_BWD = 0, K_RTRAN_FWD, K_PROJECT, K_VERFFT_FWD, K_VERFFT_BWD, K_HORFFT_FWD, K_HORFFT_BWD,
  K_PRODUCTS, K_BAND_SOLVE, K_ETD, K_SPECTRAL, K_COUNT
};
extern const char *const kernel_names[K_COUNT];
extern const char *const kernel_units[K_COUNT];  // "flop" or "byte"
struct EventPair {
  hipEvent_t a, b;
  double work;
};
// Cached band-matrix factors of one solve operator (see k_spectral.hip).
struct FactorSet {
  int op;
  double alp;
  int nrchop;
  double *f;
  int pins = 0;  // captured steps that use this set: not evicted while > 0
  double nu = 0.0;  // IHELMP sets: the del^2 coefficient (op = 100 + P)
};
struct KTimer {
  bool on = false;
  std::vector<EventPair> rec[K_COUNT];
  std::vector<hipEvent_t> spare;
};
}  // namespace mlg

namespace mlg {
struct SweepRec;
}

struct mlg_ctx {
  mlg::Grid g;
  mlg::KTimer timer;
  mlg::DevTables d;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  long n_fff = 0, n_pff = 0, n_ppp = 0, n_max = 0;
  std::vector<mlg_field *> pool;  // reusable temporaries
  std::unordered_set<mlg_field *> owned;  // fields created through the C ABI
  std::vector<mlg::FactorSet> fcache;     // band-solve factors
  int device = 0;
  int leg_split = 0;  // Legendre products in split hi/lo FP32 (mlg_set_legendre_precision)
  int split_fp64_modes = 5;  // split mode: modes m < this keep FP64 products
  // PROJECT's m = 0, k = 0 column vectors, one buffer per truncation (built on first use)
  std::vector<std::pair<int, double *>> m0vecs;
  // the adjoint's VEL2PLN vectors, one buffer per truncation (built on first use)
  std::vector<std::pair<int, double *>> adjv;
  // set while a sweep records this context's step (RTRAN backward products
  // are then deferred to the sweep's grouped launches)
  mlg::SweepRec *rec = nullptr;
  // axial sponge (RMV%SW = 2; remove.hip): TOPHAT_X over the z points, tau,
  // and the device damping per z plane for dt, dt/2 and one other step
  int rmv_sw = 0;
  std::vector<double> tophat;
  double sp_tau = 0.0;
  double *d_sponge = nullptr;
  // HYP3ADJ: NUP_R, NUP_T, NUP_X, KAPPAP_R, KAPPAP_T, KAPPAP_X and the
  // controllers' baselines and histories (CURVMO, CURVKO, SLOPMO, SLOPKO)
  double hv_coef[6] = {0, 0, 0, 0, 0, 0};
  double nupb = -1.0, kappapb = -1.0;
  double hist[4] = {0, 0, 0, 0}, phist[4] = {0, 0, 0, 0};
};

struct mlg_solver {
  mlg_ctx *ctx;
  double t = 0.0;
  int n = 0;
  mlg_field *n_psi = nullptr, *n_chi = nullptr, *n_b = nullptr;
  mlg_field *r_n_r = nullptr, *r_n_p = nullptr, *n_z = nullptr;
  mlg_field *r_pi_r = nullptr, *r_pi_p = nullptr, *pi_z = nullptr;
  double etd_e[16], etd_nl[16];
  // graph of one captured step (mlg_time_steps_graph).  The graph holds the
  // device addresses it was captured with: the temporaries it used are taken
  // out of the context's pool into gpool and the band-factor sets it reads
  // are pinned in the cache, both until the graph is dropped; gaddr records
  // the element and LN buffers of psi, chi, b at capture.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  mlg_field *gpsi = nullptr, *gchi = nullptr, *gb = nullptr;
  const void *gaddr[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<mlg_field *> gpool;
  std::vector<double *> gfactors;
  int gnrchop = -1;
  int gsplit = -1;
  hipStream_t gstream = nullptr;
  // TIME_STEPPING's REMOVE and HYP3ADJ cadence (mlg_set_march_controls)
  int rmv_sw = 0, rmv_int = 0, adj_sw = 0, adj_int = 0;
};

// Kernel-side view of the current grid (passed by value to kernels).
struct KGrid {
  int nr, nrh, nth, nx, ntchop, nxchop;
  int ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim, nxchoph;
  int nrchop, npf, nband, minc;
  int mlo, mcnt;  // Legendre product launches: the modes mlo .. mlo+mcnt-1 (all by default)
  double ell2, norm0, norm1, norm2, at1_0;
};

#define MLG_CHECK_HIP(expr)                                                                 \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      std::fprintf(stderr, "mlegs_hip: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
                   __FILE__, __LINE__);                                                     \
      return MLG_E_HIP;                                                                     \
    }                                                                                       \
  } while (0)

namespace mlg {

KGrid kgrid(const mlg_ctx *c);

// Records HIP events around one launch when the context's timer is on.
// `work` is the launch's algorithmic flops or bytes.
struct KScope {
  mlg_ctx *c;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  double work;
  KScope(mlg_ctx *c_, int id_, double work_);
  ~KScope();
};
// sum over m of NRCHOPS(m) at the current truncation
long sum_nrchops(const mlg_ctx *c);
double sum_nrchops_sq(const mlg_ctx *c);  // sum over m of NRCHOPS(m)^2

static inline int nrchops_of(int nrchop, int m) {
  int v = nrchop - m < nrchop ? nrchop - m : nrchop;
  return v > 0 ? v : 0;
}

// ---- launchers: Legendre transforms (k_legendre.hip)
int launch_rtran_bwd(mlg_ctx *c, const double2 *in, const double *ln_in, double2 *out);
// n (<= MAX_TRANSFORM_JOBS) independent fields in one launch
constexpr int MAX_TRANSFORM_JOBS = 9;
struct TransformJobs {
  const double2 *in[MAX_TRANSFORM_JOBS];
  const double *ln[MAX_TRANSFORM_JOBS];
  double2 *out[MAX_TRANSFORM_JOBS];
  unsigned dz = 0;  // VERFFT backward: bit j set = job j transforms i k times its input (d/dz)
  unsigned dm = 0;  // VERFFT backward: bit j set = job j transforms i m times its input (d/dtheta)
  double *zl[MAX_TRANSFORM_JOBS];  // VERFFT backward: LN coefficients set to zero by the launch
  int nzl = 0;
};
int launch_rtran_bwd_n(mlg_ctx *c, int n, const double2 *const *in, const double *const *ln_in, double2 *const *out);
// Grouped backward products (k_legendre.hip): any number of jobs, each one
// run's product over a range of its modes with its own tile shape (radial
// nodes x real columns), in one persistent launch.  Every shape gives the
// same bits.
enum BwdTile { BWD_TILE_64x64 = 0, BWD_TILE_32x64, BWD_TILE_64x32, BWD_TILE_COUNT };
struct BwdJob {
  KGrid g;  // the run's grid; the job covers its modes g.mlo .. g.mlo + g.mcnt - 1
  const double *pf, *lnx;
  TransformJobs jobs;
  int nj, shape;  // fields of the product, tile shape (BwdTile)
};
int bwd_job_tiles(const BwdJob &j);
// blocks of a grouped launch of `total` tiles on a device with `cus` CUs
int bwd_grouped_grid(int total, int cus);
void *bwd_grouped_kernel();  // the grouped kernel (graph kernel nodes)
// jobs[0..njobs-1] and b0[0..njobs] (first tile of each job; b0[njobs] =
// total) in device memory
int launch_rtran_bwd_grouped(hipStream_t st, const BwdJob *jobs, const int *b0, int njobs, int total, int grid);
// A step recorded for a sweep (sweep.hip): ends the capture segment at this
// backward product, keeps the product for the sweep's grouped launch and
// starts the next segment.
int sweep_defer_bwd(mlg_ctx *c, const KGrid &g, const TransformJobs &jobs, int nj);
// one AB2-CN / ETD2 step of the solver (api.hip), launches only
int step_once(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b);
// ---- vorticity removal, sponge, adaptive hyperviscosity (remove.hip)
// CALC_BOUSSI_ETD_SPONGE on the ETD update's u_z and b (no-op unless the
// sponge is on)
int etd_sponge(mlg_ctx *c, mlg_field *uz, mlg_field *b, double dt);
// the filter tables of the context's current coefficients to the device
int upload_hyperv(mlg_ctx *c);
// dz, dm: bit j set = job j transforms d/dz (i k times) or d/dtheta (i m
// times) of its input (four-step lengths, NX <= 256, only); zl[0..nzl-1]:
// LN coefficients the launch sets to zero (nzl <= MAX_TRANSFORM_JOBS)
int launch_verfft_bwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out, unsigned dz = 0,
                        unsigned dm = 0, double *const *zl = nullptr, int nzl = 0);
int launch_verfft_fwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out);
// zl: an LN set to 0; w: the Gauss weights of the sums (null: the grid's;
// the adjoint passes ones, FP64 products only)
int launch_rtran_fwd(mlg_ctx *c, const double2 *in, double2 *out, double *zl = nullptr, const double *w = nullptr);
// n (<= MAX_TRANSFORM_JOBS) fields in one launch (FP64 products only)
int launch_rtran_fwd_n(mlg_ctx *c, int n, const double2 *const *in, double2 *const *out, const double *w = nullptr);
int launch_project(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *psi, double2 *chi);
// PROJECT of spectral (FFF) velocities: RTRAN backward, VEL2PLN and the
// projection products in one pass with the composed tables; ln_mode 0: no
// VEL2PLN, else as for launch_vel2pln; the velocities' LN coefficients are
// read from ln_u.  The velocities are not changed.  PROJECT's value-at-
// infinity fix-up of psi is applied by the same launch.
int launch_project_fff(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, const double *const ln_u[3],
                       double2 *psi, double *psi_ln, double2 *chi, int ln_mode, double ln);
// Split hi/lo FP32 tables for RTRAN and PROJECT (the composed PROJECT tables
// are split when first used)
int build_split_tables(mlg_ctx *c);
// ---- FFTs (k_fft.hip)
int launch_verfft_fwd(mlg_ctx *c, const double2 *in, double2 *out);
int launch_verfft_bwd(mlg_ctx *c, const double2 *in, double2 *out);
int launch_horfft_fwd(mlg_ctx *c, double2 *a);
int launch_horfft_bwd(mlg_ctx *c, double2 *a);
// n (<= MAX_HOR_JOBS) fields in place in one launch.  Forward: the results
// of m = 0 times s0, of m >= 1 times s1.  Backward: the inputs of m = 0 times
// s0, of m >= 1 times s1, and each radial row r times wr(min(r, NR-1-r))
// when wr is given.  (s0 = s1 = 1, wr = null: HORFFT itself.)
constexpr int MAX_HOR_JOBS = 10;
struct HorJobs {
  double2 *a[MAX_HOR_JOBS];
  double s0, s1;
  const double *wr;
};
int launch_horfft_fwd_n(mlg_ctx *c, int n, double2 *const *a, double s0, double s1);
int launch_horfft_bwd_n(mlg_ctx *c, int n, double2 *const *a, double s0, double s1, const double *wr);
// ---- physical-space products (k_physical.hip)
int launch_vprod(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *ror, double2 *rop, double2 *oz);
int launch_vdot(mlg_ctx *c, const double2 *rar, const double2 *rap, const double2 *az, const double2 *rbr,
                const double2 *rbp, const double2 *bz, double2 *out, double sign);
int launch_force_products(mlg_ctx *c, const double2 *rur, const double2 *rup, const double2 *uz, double2 *ror, double2 *rop,
                          double2 *oz, const double2 *rbr, const double2 *rbp, const double2 *bz, const double2 *bphys,
                          double2 *bn, double two_omega, double n2, int linear);
// ln_mode: 0 psi LN = the computed value, 1 psi LN = ln, 2 psi LN unchanged
int launch_vel2pln(mlg_ctx *c, double2 *rup, double *psi_ln, int ln_mode, double ln);
// the same with the projections in a fixed per-device array of at most
// VEL2PLN_SCRATCH_MAX entries (no allocation); vel2pln_scratch_setup looks up
// the array's address on the current device, once per device
constexpr int VEL2PLN_SCRATCH_MAX = 1024;
constexpr int MLG_MAX_DEVICES = 64;
int vel2pln_scratch_setup(int device);
int launch_vel2pln_noalloc(mlg_ctx *c, double2 *rup, double *psi_ln, int ln_mode, double ln);
// ---- fused theta transforms + products, PFP in and out (k_fused.hip)
bool fused_products_supported(const mlg_ctx *c);
// in: rur rup uz ror rop oz rbr rbp bz [W]; out: r N_r, r N_theta, N_z, N_b;
// zl: an LN coefficient the launch sets to zero (or null)
int launch_fused_products(mlg_ctx *c, const double2 *const in[10], double2 *const out[4], double two_omega, double n2,
                          int linear, double *zl = nullptr);
// ---- spectral-space kernels (k_spectral.hip)
int launch_chopdo(mlg_ctx *c, double2 *a, int space);
constexpr int MAX_FIELD_LIST = 4;
struct FieldList {
  double2 *e[MAX_FIELD_LIST];
};
// CHOPDO of n FFF fields in one launch; fix: each field's value-at-infinity
// fix-up (launch_fix_at_infinity) in the same launch
int launch_chopdo_fff_n(mlg_ctx *c, int n, double2 *const *a, bool fix);
int launch_zero(mlg_ctx *c, double2 *a, long n);
int launch_copy(mlg_ctx *c, double2 *dst, const double2 *src, long n);
int launch_copy0_fff(mlg_ctx *c, double2 *dst, double *dst_ln, const double2 *src, const double *src_ln);
// VISC2's right-hand side WK = -CHOP(DEL2 at NRCHOP+2 of AO) - ALP*A, WK%LN = -ALP*A%LN
struct Visc2Job {
  const double2 *ao;
  const double *ln_ao;
  const double2 *a;
  const double *ln_a;
  double alp;
  double2 *wk;
  double *ln_wk;
};
struct Visc2Jobs {
  Visc2Job j[3];
};
// n (<= 3) independent right-hand sides in one launch, WK written
// row-interleaved over the systems (row n of system s = m + NTCHOPDIM k at
// wk[n * NTCHOPDIM * NXCHOPDIM + s]), as the band solve reads it
// il: WK row-interleaved over the systems (the one-thread-per-system
// substitution), else in the FFF columns (the partitioned substitution)
int launch_visc2_rhs_il(mlg_ctx *c, int n, const Visc2Jobs &jobs, int il);
// lanes per penta-diagonal system of the band substitution (0: one thread
// per system, which reads row-interleaved right-hand sides coalesced)
int band_lanes(const KGrid &g);
int launch_set_ln(mlg_ctx *c, double *dst, const double *src, double scale, double add);
int launch_zero_lns(mlg_ctx *c, double *const *ptrs, int n);  // LN = 0 for n fields, one launch per 32
int launch_delsqh(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *b, double *ln_b);
int launch_idelsqh(mlg_ctx *c, const double2 *b, double2 *a, double *ln_a);
int launch_band3(mlg_ctx *c, int which, const double2 *a, const double *ln_a, double2 *b, double *ln_b);
int launch_del2(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *b, double *ln_b);
// zl0, zl1: further LN coefficients the launch sets to zero (or null)
int launch_grad(mlg_ctx *c, const double2 *a, const double *ln_a, double2 *rr, double *ln_rr, double2 *rp, double *ln_rp,
                double2 *z, double *ln_z, double *zl0 = nullptr, double *zl1 = nullptr);
// fix_nrchop >= 1: psi and chi come straight from PROJECT and IDEL2 at the
// truncation fix_nrchop (<= NRCHOP), without IDEL2's CHOPDO and either
// value-at-infinity fix-up, which are applied to the staged columns only
// (for a psi and chi that nothing else reads)
int launch_pc2vel(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                  double2 *rur, double2 *rup, double2 *uz, double *ln_rur, double *ln_rup, double *ln_uz,
                  int fix_nrchop = 0);
// PC2VEL whose velocities only update the histories: H = H - u (hmode 4) or
// H = ha (H - u) (hmode 3), the AXPBY modes of the same names
// three FFF element arrays (null e[0]: not given)
struct VelTriple {
  const double2 *e[3];
};
// H = H - u (hmode 4) or H = ha (H - u) (hmode 3) with u = PC2VEL(psi, chi);
// H on the right is read from `base` when base.e[0] is not null
int launch_pc2vel_hist(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                       double2 *hr, double2 *hp, double2 *hz, int hmode, double ha, const VelTriple &base);
// PC2VEL and PC2VOR of one (psi, chi): outputs rur, rup, uz, ror, rop, oz
struct VelVorOut {
  double2 *e[6];
  double *ln[6];
};
int launch_pc2velvor(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                     const VelVorOut &o);
int launch_pc2vor(mlg_ctx *c, const double2 *psi, const double *ln_psi, const double2 *chi, const double *ln_chi,
                  double2 *ror, double2 *rop, double2 *oz, double *ln_ror, double *ln_rop, double *ln_oz);
int launch_fix_at_infinity(mlg_ctx *c, double2 *a);
// op: 0 IDEL2, 1 IDEL2LN, 2 IHELM
// A = solve(B) on the FFF coefficients inside the truncation (B may be A);
// the caller chops A afterwards.  t: scratch of FFF size (no other use
// during the solve)
// IHELMP (src/mod_legops.f90:1409-1592): a = (del^P + nu del^2 - alp)^-1 b
// for even P in 4..14, with its log term, value at infinity and CHOPDO; a
// may equal b
int launch_ihelmp(mlg_ctx *c, int p, const double2 *b, const double *ln_b, double2 *a, double *ln_a, double alp, double nu);
int launch_band_solve(mlg_ctx *c, int op, const double2 *b, double2 *a, double *ln_a, const double *ln_b, double alp,
                      int has_ln, double ln, double2 *t);
// n (<= 3) independent solves of one operator in one launch (field i with
// alp[i], scratch t[i]); bil: the right-hand sides are row-interleaved as
// written by launch_visc2_rhs_il
int launch_band_solve_n(mlg_ctx *c, int op, int n, const double2 *const *b, double2 *const *a, double *const *ln_a,
                        const double *const *ln_b, const double *alp, int has_ln, double ln, double2 *const *t,
                        int bil = 0);
// Adjoint of n (<= 3) op 1 (IDEL2LN) or op 2 (IHELM) solves in one launch:
// job i reads the solution's adjoint from in (FFF, chopped on the way in)
// and writes the right-hand side's adjoint to out (may equal in; CHOPDO
// included); ln_in: LNbar of the solution; ln_out (op 2): LNbar of the
// right-hand side (may equal ln_in).  op 1 includes the transpose of
// IDEL2LN's fix at infinity.  F is filled in by the launcher.
struct BandTJob {
  const double2 *in;
  double2 *out;
  const double *ln_in;
  double *ln_out;
  double alp;
  const double *F;
};
int launch_band_solve_t_n(mlg_ctx *c, int op, int n, const BandTJob *jobs);
int launch_axpby(mlg_ctx *c, int mode, double2 *y, double a, const double2 *x1, const double2 *x2, const double2 *x3, long n);
// STEP_BOUSSI_AB_CN's middle for n (<= 3) FFF fields in one launch: y0 =
// COPY0(y), y = AB2(y, x, h), h = COPY0(x), with their LN (y's AB2 only
// where ln is set)
struct Ab2Job {
  double2 *y, *y0, *h;
  const double2 *x;
  double *y_ln, *y0_ln, *h_ln;
  const double *x_ln;
  int ln;
};
struct Ab2Jobs {
  Ab2Job j[3];
};
int launch_ab2_hist(mlg_ctx *c, int n, const Ab2Jobs &jobs, double dt);
int launch_ln_axpby(mlg_ctx *c, int mode, double *y, double a, const double *x1, const double *x2, const double *x3);
// ETD update fused with the HYPERV3 filters (filter != 0), the copy of b
// before the update (b0 != null) and the history stores (save bit 0: the
// nonlinear terms, bit 1: the pressure-gradient terms); the updated b goes
// to bout (null: b itself); lnd/lns (optional): two LN copies *lnd[i] =
// *lns[i] made by the same launch (null entries skipped)
int launch_etd_sub(mlg_ctx *c, const double *e16, const double *nl16, double dt, double2 *rur, double2 *rup, double2 *uz,
                   double2 *b, const double2 *nr_, const double2 *np_, const double2 *nz_, const double2 *nb_, double2 *hpr,
                   double2 *hpp, double2 *hpz, double2 *hnr, double2 *hnp, double2 *hnz, double2 *hnb, int ab, int save,
                   double2 *b0, double2 *bout, int filter, double tf, double *const *lnd = nullptr,
                   const double *const *lns = nullptr,  // lnd, lns: 3 entries, *lnd[i] = *lns[i] (null lns[i]: 0)
                   int post_filter = 1);  // 0: the filter acts on the nonlinear terms only (the caller filters u, b)
int launch_hyp3(mlg_ctx *c, double2 *a, int diff, double tf);
int launch_diagnost(mlg_ctx *c, const double2 *psi);
int launch_energy_terms(mlg_ctx *c, const double2 *a, const double2 *b, int l2, double *out_all);
int launch_negate(mlg_ctx *c, double2 *a, long n);
int launch_scale(mlg_ctx *c, double2 *a, double s, long n);
int launch_check_top_m1(mlg_ctx *c, const double2 *a);

// ---- adjoint kernels (k_adjoint.hip)
int launch_adj_scale_pff(mlg_ctx *c, double2 *a, int mode, double s, int skip_m0);
int launch_adj_scale_pfp(mlg_ctx *c, double2 *a, double s0, double s1);
int launch_adj_products(mlg_ctx *c, double2 *const u[9], const double2 *const o[3], double2 *bn);
int launch_adj_project(mlg_ctx *c, const double2 *psib, const double2 *chib, double2 *vr, double2 *vp, double2 *vz);
int launch_adj_vel2pln(mlg_ctx *c, double2 *vp);
int launch_adj_fixinf(mlg_ctx *c, double2 *a);
int launch_adj_columns(mlg_ctx *c, int nrchop_out, const double2 *const in[9], double2 *psib, double *psib_ln, double2 *chib,
                       double *chib_ln, double2 *bb);
int launch_adj_del2t(mlg_ctx *c, int nrchop_out, const double2 *y, double s, double2 *out, double *ln_out);
int launch_adj_lincomb(mlg_ctx *c, double2 *y, double a, const double2 *x1, double b1, const double2 *x2, double b2, long n);
int launch_adj_ln_lincomb(mlg_ctx *c, double *y, double a, const double *x1, double b1, const double *x2, double b2);
// one field of the AB2-CN update's transpose (k_adj_cn)
struct AdjCnField {
  const double2 *wk;
  const double *wk_ln;
  double2 *ub, *hb, *nb;
  double *ub_ln, *hb_ln, *nb_ln;  // null: the field has no log term
  double alp;
  int cn;
};
int launch_adj_cn(mlg_ctx *c, int nrchop_out, int nf, const AdjCnField *f, double dt);

// ---- pooled temporaries and COPY0 (api.hip)
mlg_field *pool_take(mlg_ctx *c, int space);
void pool_give(mlg_ctx *c, mlg_field *f);
int copy_field(mlg_field *b, const mlg_field *a);  // COPY0: LN included, FFF copies chopped

// Temporaries of one routine.  get() hands out pool buffers; ready() checks
// them and zeroes all their LN coefficients with one launch.
struct Temps {
  mlg_ctx *c;
  std::vector<mlg_field *> v;
  std::vector<double *> pending;
  bool failed = false;
  explicit Temps(mlg_ctx *cc) : c(cc) {}
  mlg_field *get(int space = MLG_FFF) {
    mlg_field *f = pool_take(c, space);
    if (f) {
      v.push_back(f);
      pending.push_back(f->ln);
    } else {
      failed = true;
    }
    return f;
  }
  // a buffer whose LN is never read (it is copied over, or only its E is used)
  mlg_field *buffer(int space = MLG_FFF) {
    mlg_field *f = pool_take(c, space);
    if (f) v.push_back(f);
    else failed = true;
    return f;
  }
  int ready() {
    if (failed) return MLG_E_NOMEM;
    int rc = pending.empty() ? 0 : launch_zero_lns(c, pending.data(), (int)pending.size());
    pending.clear();
    return rc;
  }
  ~Temps() {
    for (auto *f : v) pool_give(c, f);
  }
};

}  // namespace mlg
