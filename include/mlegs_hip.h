/*
 * mlegs_hip.h -- C ABI of libmlegs_hip.so, the MI355X (gfx950) implementation
 * of the MLegS pseudo-spectral Boussinesq time step.
 *
 * The reference (Jinge-Wang/MLegS) has no FFI: its interface is a set of
 * Fortran modules.  Every entry point below replaces one module procedure;
 * the comment gives the reference routine (file:line) it stands in for.  A
 * Fortran iso_c_binding shim and a Python ctypes layer bind these symbols.
 *
 * Conventions
 *  - Plain C types only.  Complex values are interleaved (re, im) doubles.
 *  - Host arrays use the reference layout (column-major E(d1,d2,d3)):
 *      FFF  E(NRCHOPDIM, NTCHOPDIM, NXCHOPDIM)  spectral (n, m, k)
 *      PFF  E(NDIMR,     NTCHOPDIM, NXCHOPDIM)  (r, m, k)
 *      PFP  E(NDIMR,     NDIMTH,    NDIMX)      (r, m, z)
 *      PPP  E(NDIMR,     NDIMTH,    NDIMX)      (r, theta slot, z)
 *    with NDIMR=NR+3, NDIMTH=NTH+1, NDIMX=NX, NRCHOPDIM=min(NRCHOP+6,NR),
 *    NTCHOPDIM=min(NTCHOP,NTH), NXCHOPDIM=2*NXCHOP-1 (mod_fft.f90:222-235).
 *  - Device data stay on the GPU; every call is asynchronous on the context's
 *    stream except the ones documented as synchronising (downloads, error
 *    reads).  Field log-term coefficients (SCALAR%LN) live on the device too.
 *  - Every function returns an int status using the reference's error codes
 *    (mod_scalar3.f90:88-106): 0 ok, 1 SCALAR3, 10 NaN, 11 dealiasing,
 *    12 divergence, 13 log term, 2 TRANSFORM, 3 LEGOPERATOR, 30 deprecated,
 *    31 CALCULATION, 32 DIMMISMATCH, 4 MPI, 5 SOLVER; plus 100+ for
 *    library/runtime errors (MLG_E_*).  The reference aborts on these; the
 *    library returns them and the Fortran shim turns them into ERROR STOP.
 *  - A context and its fields are used from one host thread at a time.
 */
#ifndef MLEGS_HIP_H
#define MLEGS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Grid and physics parameters (read.input blocks DIMENSION, CHOPS, ELL,
 * ZLEN, DT, VISCOSITY, BOUSSINESQ; mod_init.f90:401-710). */
typedef struct mlg_params {
  int nr, nth, nx;            /* collocation points: r, theta slots, z   */
  int nrchop, ntchop, nxchop; /* truncations                            */
  int minc, mklink;           /* azimuthal symmetry, m-k link (0)        */
  double ell, zlen;           /* map parameter L, axial period           */
  double nu, kappa;           /* viscosity, diffusivity                  */
  double nup, kappap;         /* hyperviscosity / hyperdiffusivity       */
  int visc_sw, visc_p;        /* VISC%SW (0..3), hyperviscosity power P  */
  double omega, bv0;          /* rotation rate, Brunt-Vaisala frequency  */
  int adams;                  /* 1: AB2-CN, 0: ETD2-AB with CN           */
  int hyper_cn;               /* 1: also apply the SW=3 filter with CN   */
  double dt;
} mlg_params;

typedef struct mlg_ctx mlg_ctx;
typedef struct mlg_field mlg_field;
typedef struct mlg_solver mlg_solver;

/* space tags: mod_scalar3.f90:82-85 */
enum { MLG_PPP = 0, MLG_FFF = 1, MLG_PFP = 2, MLG_PFF = 3 };

/* library error codes (reference codes are listed above) */
enum {
  MLG_E_ARG = 100,      /* bad argument / unsupported configuration */
  MLG_E_HIP = 101,      /* HIP runtime error                        */
  MLG_E_NOMEM = 102,    /* device allocation failed                 */
  MLG_E_SPACE = 103,    /* field in the wrong space for the call    */
  MLG_E_SIZE = 104,     /* input larger than the fixed scratch of an
                           allocation-free entry point              */
  MLG_E_IO = 105        /* a file could not be opened, read or written */
};

/* ------------------------------------------------------------------ */
/* Set-up: LEGINIT / SETUP_GRID (src/mod_fft.f90:165-367).  Builds the
 * Gauss-Legendre grid, normalisations and the Pbar_n^m tables on the host
 * (extended precision) and uploads them.  Uses the current HIP device. */
int mlg_init(const mlg_params *p, mlg_ctx **ctx);
/* Parameters from a read.input file: the DIMENSION, CHOPS, ELL, ZLEN,
 * MINC, MKLINK, DT, VISCOSITY and BOUSSINESQ blocks of READIN0
 * (src/mod_init.f90:401-710, line syntax of READCOM1 :165-372), with its
 * conversions (ZLEN <= 1 is a wavenumber, NU/NUP >= 1 are Reynolds
 * numbers).  *t_limit (may be NULL) receives DT's end time.  Host only. */
int mlg_read_input(const char *path, mlg_params *p, double *t_limit);
/* Frees every device resource of the context (FINALIZE, mod_march.f90:310-353). */
int mlg_finalize(mlg_ctx *ctx);
/* out[12] = {NDIMR, NDIMTH, NDIMX, NRCHOPDIM, NTCHOPDIM, NXCHOPDIM, NRH,
 *            NXCHOPH, NRCHOP (current), NPF, NNORM, NR} */
int mlg_dims(const mlg_ctx *ctx, int *out);
/* Host copy of a set-up table: "x","w","r","lnx","pf","lognorm","norm",
 * "at0","at1","ak".  n = capacity in doubles; returns MLG_E_ARG if short. */
int mlg_get_table(const mlg_ctx *ctx, const char *name, double *out, long n, long *len);
/* Use an existing HIP stream (hipStream_t) for all work of this context;
 * NULL selects the context's own stream. */
int mlg_set_stream(mlg_ctx *ctx, void *hip_stream);
/* Blocks until all queued work of the context is done; returns the first
 * sticky device-side error code raised since the last call, then clears it. */
int mlg_synchronize(mlg_ctx *ctx);
/* Human-readable name of a status code. */
const char *mlg_error_string(int code);

/* CHOPSET (src/mod_scalar3.f90:165-211): NRCHOP += iof. */
int mlg_chopset(mlg_ctx *ctx, int iof);
int mlg_nrchop(const mlg_ctx *ctx);

/* ------------------------------------------------------------------ */
/* Fields: TYPE(SCALAR) (src/mod_scalar3.f90:41-47) with ALLOCATE/DEALLOCATE
 * (src/mod_fft.f90:603-704).  A field owns device storage for the largest
 * space plus its device log-term coefficient. */
int mlg_field_create(mlg_ctx *ctx, int space, mlg_field **f);
int mlg_field_destroy(mlg_field *f);
int mlg_field_space(const mlg_field *f);
/* Host <-> device copies in the reference layout of `space` (synchronising). */
int mlg_field_upload(mlg_field *f, const double *e, int space, double ln);
int mlg_field_download(const mlg_field *f, double *e, int *space, double *ln);
/* Raw device pointers (element storage and the LN double) for zero-copy use
 * from PyTorch or other device code on the same stream. */
int mlg_field_device_ptr(const mlg_field *f, void **e, void **ln);
/* Set the space tag and zero the field (e.g. ALLOCATE + A%E = 0). */
int mlg_field_zero(mlg_field *f, int space);
/* B = A (COPY0, src/mod_fft.f90:706-769). */
int mlg_copy(mlg_field *b, const mlg_field *a);
/* CHOPDO (src/mod_scalar3.f90:213-257). */
int mlg_chopdo(mlg_field *a);

/* ------------------------------------------------------------------ */
/* Transforms: is = -1 physical -> spectral, +1 spectral -> physical.    */
int mlg_horfft(mlg_field *a, int is);   /* HORFFT  src/mod_fft.f90:771-926   */
int mlg_verfft(mlg_field *a, int is);   /* VERFFT  src/mod_fft.f90:928-1113  */
int mlg_rtran(mlg_field *a, int is);    /* RTRAN   src/mod_fft.f90:1150-1340 */
int mlg_toff(mlg_field *a);             /* TOFF    src/mod_fft.f90:1342-1369 */
int mlg_tofp(mlg_field *a);             /* TOFP    src/mod_fft.f90:1371-1398 */
int mlg_plnfft(mlg_field *a, int is);   /* PLNFFT  src/mod_fft.f90:1115-1148 */

/* ------------------------------------------------------------------ */
/* Spectral operators (src/mod_legops.f90).  Argument order follows the
 * reference calls: (input, output). */
int mlg_delsqh(const mlg_field *a, mlg_field *b);              /* DELSQH  :76-133    */
int mlg_idelsqh(const mlg_field *b, mlg_field *a);             /* IDELSQH :136-239   */
int mlg_mulxm(const mlg_field *a, mlg_field *b);               /* MULXM   :371-428   */
int mlg_xxdx(const mlg_field *a, mlg_field *b);                /* XXDX    :680-748   */
int mlg_del2(const mlg_field *a, mlg_field *b);                /* DEL2    :751-820   */
int mlg_idel2(const mlg_field *b, mlg_field *a, int has_ln, double ln); /* IDEL2 :823-1010 */
int mlg_idel2ln(const mlg_field *b, mlg_field *a);             /* IDEL2LN :1013-1176 */
int mlg_ihelm(const mlg_field *b, mlg_field *a, double alp);   /* IHELM   :1204-1354 */
/* IHELMP :1409-1592: a = (del^P + nu del^2 - alp)^-1 b, even P in 4..14,
 * alp != 0 (else 3), with the log term at m = 0, k = 0, the value at
 * infinity of m = 0 set to zero, and CHOPDO.  VISC%SW = 2's HYPERV step. */
int mlg_ihelmp(const mlg_field *b, mlg_field *a, int p, double alp, double nu);
int mlg_grad(const mlg_field *a, mlg_field *rr, mlg_field *rp, mlg_field *z); /* GRAD :1727-1786 */
int mlg_pc2vel(const mlg_field *psi, const mlg_field *chi,
               mlg_field *rur, mlg_field *rup, mlg_field *uz); /* PC2VEL  :1789-1853 */
int mlg_pc2vor(const mlg_field *psi, const mlg_field *chi,
               mlg_field *ror, mlg_field *rop, mlg_field *oz); /* PC2VOR  :1856-1919 */
/* VPRODSUB (:1972-2133): (ror,rop,oz) <- u x omega in PPP, in place. */
int mlg_vprodsub(mlg_field *rur, mlg_field *rup, mlg_field *uz,
                 mlg_field *ror, mlg_field *rop, mlg_field *oz);
/* VDOTPSUB (:2136-2232): out <- (a . b) with the r^-2 horizontal weight. */
int mlg_vdotpsub(const mlg_field *rar, const mlg_field *rap, const mlg_field *az,
                 const mlg_field *rbr, const mlg_field *rbp, const mlg_field *bz, mlg_field *out);
int mlg_vel2pln(mlg_field *rup, mlg_field *psi, int has_ln, double ln); /* VEL2PLN :2277-2339 */
/* mlg_vel2pln without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments and results as mlg_vel2pln
 * (bitwise); its scratch (the NRCHOP projections) is a fixed array of 1024
 * doubles per GPU, so NRCHOP > 1024 returns MLG_E_SIZE.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
int mlg_vel2pln_noalloc(mlg_field *rup, mlg_field *psi, int has_ln, double ln);
/* PROJECT (:2397-2545): (r u_r, r u_theta, u_z) in PPP/PFF/FFF -> (psi, del^2 chi). */
int mlg_project(mlg_field *rur, mlg_field *rup, mlg_field *uz,
                mlg_field *psi, mlg_field *chi, int has_ln, double ln);

/* ------------------------------------------------------------------ */
/* Boussinesq terms (src/mod_boussinesq.f90). */
/* CALC_BOUSSI_FORCE :317-410 -> (psin, chin, bn) in FFF. */
int mlg_force(const mlg_field *psi, const mlg_field *chi, const mlg_field *b,
              mlg_field *psin, mlg_field *chin, mlg_field *bn);
/* CALC_BOUSSI_NONLIN :235-314; on exit ror/rop/oz hold (rN_r, rN_theta, N_z), bn -u.grad b, all FFF. */
int mlg_nonlin(const mlg_field *b, mlg_field *rur, mlg_field *rup, mlg_field *uz,
               mlg_field *bn, mlg_field *ror, mlg_field *rop, mlg_field *oz);
/* CALC_BOUSSI_ENERGY :138-232 (file output omitted): ke/pe/dke, each
 * NTCHOP x NXCHOP column-major (synchronising). */
int mlg_energy(const mlg_field *psi, const mlg_field *chi, const mlg_field *b,
               double *ke, double *pe, double *dke);
/* CALC_BOUSSI_ETD_OP (src/mod_march.f90:906-945): real parts of E and Phi, 4x4 column-major. */
int mlg_etd_op(const mlg_ctx *ctx, double dt, double *e16, double *nl16);

/* ------------------------------------------------------------------ */
/* Time march: SOLVER_T (src/mod_march.f90:17-39). */
int mlg_solver_create(mlg_ctx *ctx, mlg_solver **s);
int mlg_solver_destroy(mlg_solver *s);
/* INITIALIZE (:171-240): Richardson start (AB: RICH_BOUSSI_FE_BE :535-606;
 * ETD: RICH_BOUSSI_ETDFE_BE :756-808); advances one step. */
int mlg_initialize(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b);
/* TIME_STEPPING (:243-307): one STEP_BOUSSI_AB_CN (:609-661) or
 * STEP_BOUSSI_ETDAB_CN (:811-901) plus the per-step DIAGNOST checks
 * (:1726-1791).  Synchronises to return the check result. */
int mlg_time_step(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b);
/* nsteps TIME_STEPPING calls queued back to back; the checks run on the
 * device every step and the first failure is returned at the end. */
int mlg_time_steps(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps);
/* The same stepping through a HIP graph: the first call on (psi, chi, b)
 * runs one step eagerly, captures the next one and replays the captured
 * step for the rest; later calls on the same fields only replay.  No
 * counterpart in the reference (it replaces the per-step launch sequence of
 * SOLVER_T%TIME_STEPPING, src/mod_march.f90:243-307).  Results equal
 * mlg_time_steps bit for bit. */
int mlg_time_steps_graph(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps);
/* DIAGNOST (src/mod_march.f90:1726-1791), its per-step checks on given
 * fields: returns 11 (dealiasing: psi non-zero beyond NRCHOPS(m) at m = 0,
 * 1) or 12 (divergence: |psi| > 100 at n < 5, 1 <= m < 6, k < 5), else 0.
 * Synchronising; an earlier unread device-side error is returned first. */
int mlg_diagnost(const mlg_field *psi, const mlg_field *chi);
double mlg_solver_time(const mlg_solver *s);
int mlg_solver_nstep(const mlg_solver *s);

/* ------------------------------------------------------------------ */
/* Precision of the Legendre products (RTRAN, PROJECT).  No counterpart in
 * the reference, whose products are FP64 DGEMMs (src/mod_eig.f90:502-656
 * behind src/mod_fft.f90:1208-1307 and src/mod_legops.f90:2496-2648).
 * MLG_LEG_SPLIT32: table and data operands split into hi/lo FP32 on FP32
 * MFMA, exact hi x hi sums within 16-deep K chunks, FP32 cross terms, the
 * chunks combined in FP64 (relative error ~1e-10 of the operands' scales).
 * The lowest azimuthal modes keep FP64 products (m < 5 unless set with
 * mlg_set_split_fp64_modes, at least 2): their projection tables divide by
 * (1 - x^2) and their products cancel in psi and chi.  Band solves, log
 * terms, FFTs and diagnostics stay FP64.  The first switch to SPLIT32 builds
 * the split tables (synchronous allocation). */
enum { MLG_LEG_FP64 = 0, MLG_LEG_SPLIT32 = 1 };
int mlg_set_legendre_precision(mlg_ctx *ctx, int mode);
int mlg_legendre_precision(const mlg_ctx *ctx);
int mlg_set_split_fp64_modes(mlg_ctx *ctx, int nmodes);

/* ------------------------------------------------------------------ */
/* Initial conditions: QVORTEX (src/mod_layout.f90:449-610) for one centred
 * Gaussian vortex of strength q (omega_z = 2 q exp(-r^2)), chi = 0. */
int mlg_qvortex(mlg_ctx *ctx, double q, mlg_field *psi, mlg_field *chi);

/* ------------------------------------------------------------------ */
/* Field files and run settings (the drivers' I/O). */
/* MPISAVEX (src/mod_fft.f90:2304-2377): FN.info (grid, space, LN records)
 * and FN (raw complex128, one process block).  Synchronising. */
int mlg_field_save(const mlg_field *f, const char *fn);
/* MPILOADX (src/mod_fft.f90:2510-2642): fills f from FN.info + FN written
 * by a run on n1 x n2 processes (1, 1 for files of this library); a grid
 * other than the file's returns 32 (DIMMISMATCH), a bad file MLG_E_IO. */
int mlg_field_load(mlg_field *f, const char *fn, int n1, int n2);
/* The m = 0 columns E(:,1,k), k < nk, in the field's current space, as
 * INSPECT (src/mod_march.f90:3743-3775) prints them; *nrad = first extent.
 * out = NULL only reports *nrad.  Synchronising. */
int mlg_field_m0_columns(const mlg_field *f, int nk, double *out, int *nrad);

/* The run settings of read.input that the drivers use besides mlg_params
 * (READIN0, src/mod_init.f90:401-710): file names, output schedule,
 * diagnostics cadence, vortex removal and hyperviscosity adjustment
 * switches, monitored modes.  Names are blank-padded, not terminated. */
#define MLG_NAME_LEN 72
#define MLG_MAX_OUTPUTS 512
#define MLG_MAX_MONITOR 32
typedef struct mlg_run_settings {
  char savedir[MLG_NAME_LEN], psi0[MLG_NAME_LEN], chi0[MLG_NAME_LEN], b0[MLG_NAME_LEN];
  char psii[MLG_NAME_LEN], chii[MLG_NAME_LEN], bi[MLG_NAME_LEN];
  int it1, it2;          /* DIAGNOST (defaults 40, 40)            */
  int rmv_sw, rmv_int;   /* REMOVE (defaults 0, 10000)            */
  int adj_sw, adj_int;   /* NUP_ADJUST (default switch 0)         */
  int noutputs;          /* OUTPUTS: snapshot times and names     */
  double out_t[MLG_MAX_OUTPUTS];
  char out_psi[MLG_MAX_OUTPUTS][MLG_NAME_LEN], out_chi[MLG_MAX_OUTPUTS][MLG_NAME_LEN];
  char out_b[MLG_MAX_OUTPUTS][MLG_NAME_LEN];
  int nmonitor;          /* MONITOR_MK: (m, k) pairs, start, interval */
  int monitor_mk[MLG_MAX_MONITOR][2];
  int monitor_start, monitor_interval;
} mlg_run_settings;
int mlg_read_run_settings(const char *path, mlg_run_settings *s);

/* ------------------------------------------------------------------ */
/* Discrete adjoint of the AB2-CN march (no counterpart in the reference,
 * which has no adjoint; for optimal-perturbation and sensitivity callers).
 * Exact transposes, for the inner product sum Re(conj(a) b) over stored
 * coefficients plus a%LN b%LN (psi, chi), of what mlg_initialize and
 * mlg_time_step compute with ADAMS = 1, VISC%SW = 0 or 1, no
 * hyperviscosity and FP64 Legendre products (else MLG_E_ARG).  All fields
 * are FFF; adjoint fields are ordinary fields.  States are passed in, not
 * kept: a caller that checkpoints recomputes them.
 *   mlg_force_adjoint: *_bar += CALC_BOUSSI_FORCE'(psi, chi, b)^T (psin_bar,
 *     chin_bar, bn_bar).
 *   mlg_step_adjoint: (psi, chi, b) the state before one STEP_BOUSSI_AB_CN;
 *     in: *_bar the adjoint of the state after it, h*_bar of the history
 *     (N_psi, N_chi, N_b) after it; out: the same before it.
 *   mlg_start_adjoint: the same for RICH_BOUSSI_FE_BE (mlg_initialize) from
 *     the initial state; the history adjoint comes out zero.
 *   mlg_ke_gradient: d/d(psi, chi) of the sum over m >= mmin (mmin = 1) and
 *     all k of CALC_BOUSSI_ENERGY's KE(m, k), written to psi_bar, chi_bar.
 *   mlg_solver_history: the solver's own history fields. */
int mlg_force_adjoint(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, const mlg_field *psin_bar,
                      const mlg_field *chin_bar, const mlg_field *bn_bar, mlg_field *psi_bar, mlg_field *chi_bar,
                      mlg_field *b_bar);
int mlg_step_adjoint(mlg_solver *s, const mlg_field *psi, const mlg_field *chi, const mlg_field *b, mlg_field *psi_bar,
                     mlg_field *chi_bar, mlg_field *b_bar, mlg_field *hpsi_bar, mlg_field *hchi_bar, mlg_field *hb_bar);
int mlg_start_adjoint(mlg_solver *s, const mlg_field *psi, const mlg_field *chi, const mlg_field *b, mlg_field *psi_bar,
                      mlg_field *chi_bar, mlg_field *b_bar, mlg_field *hpsi_bar, mlg_field *hchi_bar, mlg_field *hb_bar);
int mlg_ke_gradient(const mlg_field *psi, const mlg_field *chi, int mmin, mlg_field *psi_bar, mlg_field *chi_bar);
int mlg_solver_history(mlg_solver *s, mlg_field **hpsi, mlg_field **hchi, mlg_field **hb);
/* Transposes of single operators (op codes in adjoint.hip): 0 TOFP, 1 TOFF,
 * 2 PROJECT from PPP, 3 IDEL2LN, 4 IHELM(alp), 5 DEL2, 6 PC2VEL+PC2VOR+GRAD,
 * 7 CALC_BOUSSI_FORCE's pointwise terms.  No reference counterpart. */
int mlg_adjoint_op(int op, mlg_field *const *f, int nf, double alp);

/* ------------------------------------------------------------------ */
/* DIAGNOST's every-IT1 section (src/mod_march.f90:1797-1961) with CALCXY
 * (:1967-2047), at NRCHOP+3 as the reference forms it.  out[0] VALVOR
 * (INTEG of omega_z/(1-x)^2), [1] VALZMO (of u_z/(1-x)^2), [2] VALENE,
 * [3] VALHEL, [4] VALANG (-INTEG_P(XXDX(psi))), [5] |PSIAT1| at k = 0 and
 * [6] max - min of |PSIAT1| over k (CALCAT1), [7], [8] the same for chi,
 * [9] <X>, [10] <Y>, [11] <R> (the enstrophy-weighted centre and radius of
 * CALCXY).  The printing, coresize.dat and the ADV history stay with the
 * caller. */
int mlg_diagnost_it1(const mlg_field *psi, const mlg_field *chi, double out[12]);

/* ------------------------------------------------------------------ */
/* Vorticity removal, axial sponge and adaptive hyperviscosity
 * (src/mod_march.f90), with TIME_STEPPING's cadence.
 *   mlg_set_march_controls: RMV%SW, RMV%INT, VISC%ADJSW, VISC%ADJINT
 *     (READIN0's REMOVE and NUP_ADJUST blocks).  Every RMV%INT steps
 *     (RMV%SW != 0) mlg_time_step(s) and mlg_time_steps_graph apply REMOVE,
 *     the Richardson restart and DIAGNOST; every ADJINT steps (ADJSW != 0)
 *     HYP3ADJ and HYP3SET (TIME_STEPPING :264-287).  RMV%SW = 2 also turns on
 *     the axial sponge of the ETD steps (CALC_BOUSSI_ETD_SPONGE :1436-1469;
 *     SPNGINIT :2292-2323): set it before mlg_initialize, as INITIALIZE
 *     makes SPNGINIT before its Richardson start.
 *   mlg_remove: REMOVE (:2503-2551), b may be NULL.
 *   mlg_ringcut: RINGCUT (:2403-2445) on a PFF field, ns0/ns 1-based.
 *   mlg_sponge: SPONGE (:2380-2420) of psi, chi (and b unless NULL) over dt;
 *     a no-op unless RMV%SW = 2.
 *   mlg_enemon: ENEMON (:2581-2617) at the current truncation: em[NTCHOP],
 *     ek[NXCHOP].
 *   mlg_hyp3adj: HYP3ADJ (:3000-3182) then HYP3SET of the new NUP_R/T/X and
 *     KAPPAP_R/T/X; b may be NULL (no KAPPAP update).
 *   mlg_hyperv_coeffs / mlg_set_hyperv_coeffs: NUP_R, NUP_T, NUP_X,
 *     KAPPAP_R, KAPPAP_T, KAPPAP_X (set: HYP3SET of the given values).
 *   mlg_sponge_table: TOPHAT_X over the NDIMX z points (zeros when off). */
int mlg_set_march_controls(mlg_solver *s, int rmv_sw, int rmv_int, int adj_sw, int adj_int);
int mlg_remove(mlg_field *psi, mlg_field *chi, mlg_field *b);
int mlg_ringcut(mlg_field *wk, int ns0, int ns, int divxm2);
int mlg_sponge(mlg_field *psi, mlg_field *chi, mlg_field *b, double dt);
int mlg_enemon(const mlg_field *psi, const mlg_field *chi, double *em, double *ek);
int mlg_hyp3adj(mlg_ctx *ctx, const mlg_field *psi, const mlg_field *chi, const mlg_field *b);
int mlg_hyperv_coeffs(const mlg_ctx *ctx, double out[6]);
int mlg_set_hyperv_coeffs(mlg_ctx *ctx, const double in[6]);
int mlg_sponge_table(const mlg_ctx *ctx, double *tophat, int n);

/* ------------------------------------------------------------------ */
/* Batched sweep of independent runs (no counterpart in the reference, which
 * runs one case per job through SOLVER_T%TIME_STEPPING,
 * src/mod_march.f90:244-300; for parameter-sweep callers).  Each run is a
 * solver with its own context (grid, parameters) and its state psi, chi, b
 * (FFF), already started (mlg_initialize).  mlg_sweep_steps advances every
 * run by nsteps steps; every run's results are bitwise those of
 * mlg_time_steps on that run alone.  The first call takes one eager step of
 * every run, then records each run's step into one hipGraph for all runs:
 * the runs' segments are independent; with MLG_SWEEP_GROUPED, the RTRAN
 * backward products at the same place of every run's step are one grouped
 * persistent launch over (run, mode bucket) jobs, with tile shapes chosen
 * by the tuner (MLG_SWEEP_TUNE) or given as (shape + 1) << 4 (shapes as
 * mlg_tune_rtran_bwd's).  MLG_SWEEP_STREAMS: instead of one graph, each
 * run's segments are graphs of their own launched on the run's stream (runs
 * on different streams run at once), the grouped launches on the sweep's
 * stream, joined to the runs by events.  MLG_SWEEP_BY_GRID (streams mode
 * only): one grouped launch per place and class of runs of one grid (NR,
 * NTH, NX and truncations), so that a launch waits only for runs of equal
 * cost.
 * Grouping needs FP64 Legendre
 * products.  The
 * graph is recorded again when a run's field buffers or truncation change.
 * Forced shapes take bits 4..6 of flags.
 * The sweep owns the temporaries its graph uses until destroyed.
 *
 * The runs step without TIME_STEPPING's REMOVE / HYP3ADJ events: a solver
 * with an active cadence (mlg_set_march_controls) is rejected.
 * A run whose per-step DIAGNOST check fails keeps stepping with the others
 * (runs are independent jobs): mlg_sweep_steps returns the code of the
 * lowest-numbered run that has one (after all steps), mlg_sweep_codes each
 * run's first code since the sweep was created (0: none).  Library and HIP
 * errors (MLG_E_*) are returned at once.
 *
 * mlg_sweep_info: out[0] grouped positions per step, [1] graph nodes,
 * [2] grouped jobs per step, [3..5] jobs per tile shape, [6] grouped
 * launches per step, [7] runs.
 *
 * mlg_tune_rtran_bwd: for each mode bucket of the context's current
 * truncation (by NRCHOPS(m): 0: > 128, 1: 65..128, 2: 33..64, 3: <= 32),
 * times a backward product of nfields fields over the bucket's modes with
 * each tile shape (0: 64 nodes x 64 real columns, 1: 32x64, 2: 64x32) on
 * scratch buffers: us[4 bucket + shape] in microseconds (median of 5
 * launches; us[4 bucket + 3] = 0), shape[bucket] the fastest (-1: no
 * modes).  The choice is kept per grid and used by later sweeps. */
typedef struct mlg_sweep mlg_sweep;
#define MLG_SWEEP_GROUPED 1
#define MLG_SWEEP_TUNE 2
#define MLG_SWEEP_STREAMS 4
#define MLG_SWEEP_BY_GRID 8
int mlg_sweep_create(int nruns, mlg_solver *const *solvers, mlg_field *const *psi, mlg_field *const *chi,
                     mlg_field *const *b, int flags, mlg_sweep **out);
int mlg_sweep_steps(mlg_sweep *sw, int nsteps);
int mlg_sweep_info(const mlg_sweep *sw, int out[8]);
int mlg_sweep_codes(const mlg_sweep *sw, int *codes);
/* algorithmic GFLOP of one step's grouped launches: 2 NR NXCHOPDIM x the
 * sum over the job's modes of NRCHOPS(m), per field, summed over jobs */
int mlg_sweep_work(const mlg_sweep *sw, double *gflop_per_step);
int mlg_sweep_destroy(mlg_sweep *sw);
int mlg_tune_rtran_bwd(mlg_ctx *ctx, int nfields, int shape[4], double us[16]);

/* ------------------------------------------------------------------ */
/* Kernel event timer.  Instrumentation with no counterpart in the
 * reference (which times whole steps with CPU_TIME).  mlg_timer_enable(1)
 * clears and starts recording HIP events around every launch of the timed
 * kernel classes on the context's stream; (0) stops.  mlg_timer_read
 * synchronises and returns, for one class ("rtran_bwd", "rtran_fwd",
 * "project", "verfft_fwd", "verfft_bwd", "horfft_fwd", "horfft_bwd",
 * "products", "band_solve", "etd"), the summed event time in ms, the launch
 * count and the summed algorithmic work (flops for the Legendre classes,
 * HBM bytes read+written once for the others; *unit_is_bytes tells which). */
int mlg_timer_enable(mlg_ctx *ctx, int on);
int mlg_timer_read(mlg_ctx *ctx, const char *kernel, double *ms_total, long *launches,
                   double *work_total, int *unit_is_bytes);

#ifdef __cplusplus
}
#endif
#endif
