/*
 * mlegs_oracle.h -- CPU FP64 restatement of the MLegS Boussinesq time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library (libmlegs_hip.so).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in mlegs_amd/
 * links, imports or calls it.
 *
 * Every routine follows a reference routine (file:line given at each
 * definition in mlegs_oracle.c).  Arrays use the reference's column-major
 * layout E(dim1,dim2,dim3), 0-based here:
 *   FFF  E(NRCHOPDIM, NTCHOPDIM, NXCHOPDIM)   (n, m, k)
 *   PFF  E(NDIMR,     NTCHOPDIM, NXCHOPDIM)   (r, m, k)
 *   PFP  E(NDIMR,     NDIMTH,    NDIMX)       (r, m, z)
 *   PPP  E(NDIMR,     NDIMTH,    NDIMX)       (r, theta-slot, z)
 */
#ifndef MLEGS_ORACLE_H
#define MLEGS_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as mlg_params in include/mlegs_hip.h. */
typedef struct og_params {
  int nr, nth, nx;
  int nrchop, ntchop, nxchop;
  int minc, mklink;
  double ell, zlen;
  double nu, kappa, nup, kappap;
  int visc_sw, visc_p;
  double omega, bv0;
  int adams;
  int hyper_cn;
  double dt;
} og_params;

typedef struct og_grid og_grid;
typedef struct og_solver og_solver;

typedef struct og_field {
  double complex *e;   /* sized for the largest space (PPP) */
  double ln;
  int space;
  long len;            /* elements allocated */
} og_field;

enum { OG_PPP = 0, OG_FFF = 1, OG_PFP = 2, OG_PFF = 3 };

/* grid / tables */
og_grid *og_init(const og_params *p);
void og_free(og_grid *g);
void og_dims(const og_grid *g, int *out12);   /* ndimr ndimth ndimx nrchopdim ntchopdim nxchopdim nrh nxchoph nrchop ... */
const double *og_table(const og_grid *g, const char *name, int *n);
void og_chopset(og_grid *g, int iof);
int og_nrchop(const og_grid *g);

/* fields */
og_field *og_field_new(og_grid *g, int space);
void og_field_free(og_field *f);
void og_pool_drain(void);
long og_field_size(const og_grid *g, int space);
void og_chopdo(const og_grid *g, og_field *a);
void og_copy(const og_grid *g, og_field *b, const og_field *a);

/* transforms */
int og_horfft(const og_grid *g, og_field *a, int is);
int og_verfft(const og_grid *g, og_field *a, int is);
int og_rtran(const og_grid *g, og_field *a, int is);
int og_toff(const og_grid *g, og_field *a);
int og_tofp(const og_grid *g, og_field *a);

/* spectral operators */
int og_delsqh(const og_grid *g, const og_field *a, og_field *b);
int og_idelsqh(const og_grid *g, const og_field *b, og_field *a);
int og_mulxm(const og_grid *g, const og_field *a, og_field *b);
int og_xxdx(const og_grid *g, const og_field *a, og_field *b);
int og_del2(const og_grid *g, const og_field *a, og_field *b);
int og_idel2(const og_grid *g, const og_field *b, og_field *a, int has_ln, double ln);
int og_idel2ln(const og_grid *g, const og_field *b, og_field *a);
int og_ihelm(const og_grid *g, const og_field *b, og_field *a, double alp);
int og_ihelmp(const og_grid *g, int p, const og_field *b, og_field *a, double alp, double nu);
int og_helmp(const og_grid *g, int p, const og_field *a, og_field *b, double alp, double nu);
int og_grad(const og_grid *g, const og_field *a, og_field *rr, og_field *rp, og_field *z);
int og_pc2vel(const og_grid *g, const og_field *psi, const og_field *chi,
              og_field *rur, og_field *rup, og_field *uz);
int og_pc2vor(const og_grid *g, const og_field *psi, const og_field *chi,
              og_field *ror, og_field *rop, og_field *oz);
int og_vprodsub(const og_grid *g, og_field *rur, og_field *rup, og_field *uz,
                og_field *ror, og_field *rop, og_field *oz);
int og_vdotpsub(const og_grid *g, const og_field *rar, const og_field *rap, const og_field *az,
                const og_field *rbr, const og_field *rbp, const og_field *bz, og_field *out);
int og_vel2pln(const og_grid *g, og_field *rup, og_field *psi, int has_ln, double ln);
int og_project(const og_grid *g, og_field *rur, og_field *rup, og_field *uz,
               og_field *psi, og_field *chi, int has_ln, double ln);

/* Boussinesq physics */
int og_force(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b,
             og_field *psin, og_field *chin, og_field *bn);
int og_nonlin(og_grid *g, const og_field *b, og_field *rur, og_field *rup, og_field *uz,
              og_field *bn, og_field *ror, og_field *rop, og_field *oz);
void og_etd_op(const og_grid *g, double dt, double *e16, double *nl16); /* real parts, col-major 4x4 */
int og_energy(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b,
              double *ke, double *pe, double *dke); /* each NTCHOP x NXCHOP col-major */

/* time march */
og_solver *og_solver_new(og_grid *g);
void og_solver_free(og_solver *s);
int og_initialize(og_solver *s, og_field *psi, og_field *chi, og_field *b);
int og_solver_prepare(og_solver *s);
int og_time_step(og_solver *s, og_field *psi, og_field *chi, og_field *b);
int og_diagnost(const og_grid *g, const og_field *psi, const og_field *chi);
/* DIAGNOST's every-IT1 quantities (see mlegs_oracle.c): out[12] */
int og_diagnost_it1(og_grid *g, const og_field *psi, const og_field *chi, double *out);
double og_time(const og_solver *s);
int og_nstep(const og_solver *s);

/* vorticity removal, axial sponge, adaptive hyperviscosity (mod_march.f90) */
void og_set_sponge(og_grid *g, int rmv_sw);                      /* SPNGINIT */
int og_sponge(og_grid *g, og_field *psi, og_field *chi, og_field *b, double dt); /* SPONGE, b may be NULL */
int og_ringcut(og_grid *g, og_field *wk, int ns0, int ns, int divxm2); /* PFF; 1-based ns0, ns */
int og_smooth(double *v, int ni, int passes);                        /* SMOOTH on interleaved complex (tests) */
int og_remove(og_grid *g, og_field *psi, og_field *chi, og_field *b);   /* b may be NULL */
int og_enemon(og_grid *g, const og_field *psi, const og_field *chi, double *em, double *ek);
double og_specurv(const double *espec, int n, double ratio);
double og_speslop(const double *espec, int n, double ratio);
void og_pdajust(double *coeff, double cur, double prev, double base, const double *params, int has_thr, double thr);
int og_hyp3adj(og_grid *g, const og_field *psi, const og_field *chi, const og_field *b); /* + HYP3SET */
void og_hyperv_coeffs(const og_grid *g, double *out);       /* NUP_R,T,X, KAPPAP_R,T,X */
void og_set_hyperv_coeffs(og_grid *g, const double *in);
void og_set_march_controls(og_solver *s, int rmv_sw, int rmv_int, int adj_sw, int adj_int);

/* set-up helpers */
int og_qvortex(og_grid *g, double q, og_field *psi, og_field *chi);
void og_set_threads(int n);
void og_set_sum_order(int reverse);
void og_set_mutation(int k);   /* test-only planted mistakes, see mlegs_oracle.c */

#ifdef __cplusplus
}
#endif
#endif
