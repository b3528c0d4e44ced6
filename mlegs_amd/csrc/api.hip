// C ABI of libmlegs_hip.so: context, fields and the host-side control flow of
// the Boussinesq step.  Kernels live in k_*.hip; this file sequences them on
// the context's stream, mirroring the reference's call structure and CHOPSET
// (truncation) state so results match the reference routine by routine.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>

#include "mlegs_internal.h"

using namespace mlg;

#define RET(x)                 \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

namespace mlg {
KGrid kgrid(const mlg_ctx *c) {
  const Grid &G = c->g;
  KGrid k;
  k.nr = G.nr;
  k.nrh = G.nrh;
  k.nth = G.nth;
  k.nx = G.nx;
  k.ntchop = G.ntchop;
  k.nxchop = G.nxchop;
  k.ndimr = G.ndimr;
  k.ndimth = G.ndimth;
  k.ndimx = G.ndimx;
  k.nrchopdim = G.nrchopdim;
  k.ntchopdim = G.ntchopdim;
  k.nxchopdim = G.nxchopdim;
  k.nxchoph = G.nxchoph;
  k.nrchop = G.nrchop;
  k.npf = G.npf;
  k.nband = G.nband;
  k.minc = G.p.minc;
  k.mlo = 0;
  k.mcnt = G.ntchopdim;
  k.ell2 = G.ell2;
  k.norm0 = G.norm[0];
  k.norm1 = G.norm[1];
  k.norm2 = G.norm[2];
  k.at1_0 = G.at1[0];
  return k;
}

const char *const kernel_names[K_COUNT] = {"rtran_bwd", "rtran_fwd", "project", "verfft_fwd", "verfft_bwd", "horfft_fwd",
                                           "horfft_bwd", "products", "band_solve", "etd", "spectral"};
const char *const kernel_units[K_COUNT] = {"flop", "flop", "flop", "byte", "byte", "byte",
                                           "byte", "byte", "byte", "byte", "byte"};

static hipEvent_t take_event(mlg_ctx *c) {
  hipEvent_t e = nullptr;
  if (!c->timer.spare.empty()) {
    e = c->timer.spare.back();
    c->timer.spare.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

KScope::KScope(mlg_ctx *c_, int id_, double work_) : c(c_), id(id_), work(work_) {
  if (!c->timer.on) return;
  a = take_event(c);
  b = take_event(c);
  if (a) (void)hipEventRecord(a, c->stream);
}
KScope::~KScope() {
  if (!a || !b) return;
  (void)hipEventRecord(b, c->stream);
  c->timer.rec[id].push_back(EventPair{a, b, work});
}

long sum_nrchops(const mlg_ctx *c) {
  long s = 0;
  for (int mm = 0; mm < c->g.ntchopdim; mm++) s += nrchops_of(c->g.nrchop, mm * c->g.p.minc);
  return s;
}
double sum_nrchops_sq(const mlg_ctx *c) {
  double s = 0;
  for (int mm = 0; mm < c->g.ntchopdim; mm++) {
    const double v = nrchops_of(c->g.nrchop, mm * c->g.p.minc);
    s += v * v;
  }
  return s;
}
}  // namespace mlg

static void timer_clear(mlg_ctx *c) {
  for (auto &v : c->timer.rec) {
    for (auto &p : v) {
      c->timer.spare.push_back(p.a);
      c->timer.spare.push_back(p.b);
    }
    v.clear();
  }
}

extern "C" int mlg_timer_enable(mlg_ctx *c, int on) {
  if (!c) return MLG_E_ARG;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (on) timer_clear(c);
  c->timer.on = on != 0;
  return 0;
}

extern "C" int mlg_timer_read(mlg_ctx *c, const char *kernel, double *ms_total, long *launches, double *work_total,
                              int *unit_is_bytes) {
  if (!c || !kernel) return MLG_E_ARG;
  int id = -1;
  for (int k = 0; k < K_COUNT; k++)
    if (!std::strcmp(kernel, kernel_names[k])) id = k;
  if (id < 0) return MLG_E_ARG;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  double ms = 0.0, work = 0.0;
  for (auto &p : c->timer.rec[id]) {
    float t = 0.0f;
    MLG_CHECK_HIP(hipEventElapsedTime(&t, p.a, p.b));
    ms += t;
    work += p.work;
  }
  if (ms_total) *ms_total = ms;
  if (launches) *launches = (long)c->timer.rec[id].size();
  if (work_total) *work_total = work;
  if (unit_is_bytes) *unit_is_bytes = std::strcmp(kernel_units[id], "byte") == 0;
  return 0;
}

static long space_count(const mlg_ctx *c, int space) {
  if (space == MLG_FFF) return c->n_fff;
  if (space == MLG_PFF) return c->n_pff;
  return c->n_ppp;
}

template <class T>
static int upload(T **dst, const std::vector<T> &src) {
  if (src.empty()) {
    *dst = nullptr;
    return 0;
  }
  MLG_CHECK_HIP(hipMalloc((void **)dst, sizeof(T) * src.size()));
  MLG_CHECK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return 0;
}

static bool pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

extern "C" const char *mlg_error_string(int code) {
  switch (code) {
    case 0: return "ok";
    case 1: return "SCALAR3";
    case 10: return "SCALAR3_ISNAN";
    case 11: return "SCALAR3_DEALIASING";
    case 12: return "SCALAR3_DIVERGENCE";
    case 13: return "SCALAR3_LOGTERM";
    case 2: return "TRANSFORM";
    case 3: return "LEGOPERATOR";
    case 30: return "DEPRECATED";
    case 31: return "CALCULATION";
    case 32: return "DIMMISMATCH";
    case 4: return "MPIERROR";
    case 5: return "SOLVER";
    case MLG_E_ARG: return "invalid argument or unsupported configuration";
    case MLG_E_HIP: return "HIP runtime error";
    case MLG_E_NOMEM: return "device allocation failed";
    case MLG_E_SPACE: return "field is in the wrong space for this call";
    case MLG_E_SIZE: return "size exceeds the fixed scratch of an allocation-free entry point";
    case MLG_E_IO: return "file could not be opened, read or written";
    default: return "unknown";
  }
}

// ------------------------------------------------------------------ context
extern "C" int mlg_init(const mlg_params *p, mlg_ctx **out) {
  if (!p || !out) return MLG_E_ARG;
  *out = nullptr;
  // supported: even NR, NTH and NX from 2 to 4096 (powers of two take the
  // FFT kernels, other lengths direct DFTs), the reference's chop limits
  // (mod_fft.f90:238-244, mod_scalar3.f90:183-190)
  if (p->nr < 4 || p->nr % 2 || p->nth < 2 || p->nx < 2) return MLG_E_ARG;
  if (2 * p->nxchop > p->nx || p->nxchop < 1 || p->ntchop < 1 || p->nrchop < 1) return MLG_E_ARG;
  if (p->nth > 4096 || p->nx > 4096) return MLG_E_ARG;
  if (p->minc != 1 && p->minc != 0) return MLG_E_ARG;  // the path uses MINC = 1
  if (p->visc_sw < 0 || p->visc_sw > 3) return MLG_E_ARG;  // VISC%SW: 0 none, 1 CN, 2 IHELMP, 3 filter
  if (p->mklink != 0) return MLG_E_ARG;
  mlg_ctx *c = new mlg_ctx();
  int rc = build_grid(*p, c->g);
  if (rc) {
    delete c;
    return rc;
  }
  Grid &G = c->g;
  // HYP3INIT (mod_march.f90:3463-3480): every direction at NUP and KAPPAP
  for (int i = 0; i < 3; i++) {
    c->hv_coef[i] = p->nup;
    c->hv_coef[3 + i] = p->kappap;
  }
  (void)hipGetDevice(&c->device);
  // the allocation-free entry points' fixed scratch: its address on this device
  rc = vel2pln_scratch_setup(c->device);
  if (rc) {
    delete c;
    return rc;
  }
  c->n_fff = (long)G.nrchopdim * G.ntchopdim * G.nxchopdim;
  c->n_pff = (long)G.ndimr * G.ntchopdim * G.nxchopdim;
  c->n_ppp = (long)G.ndimr * G.ndimth * G.ndimx;
  c->n_max = std::max(c->n_fff, std::max(c->n_pff, c->n_ppp));
  DevTables &d = c->d;
  RET(upload(&d.pf, G.pf));
  RET(upload(&d.vdt, G.vdt));
  RET(upload(&d.w, G.w));
  RET(upload(&d.ones, std::vector<double>(G.nr, 1.0)));
  RET(upload(&d.r, G.r));
  RET(upload(&d.x, G.x));
  RET(upload(&d.lnx, G.lnx));
  RET(upload(&d.at1, G.at1));
  RET(upload(&d.ak, G.ak));
  RET(upload(&d.norm, G.norm));
  RET(upload(&d.lognorm, G.lognorm));
  RET(upload(&d.bxm, G.bxm));
  RET(upload(&d.bxxdx, G.bxxdx));
  RET(upload(&d.bdel2h, G.bdel2h));
  // r^2 for the products (the kernels divide by r_i^2)
  {
    std::vector<double> r2(G.nr);
    for (int i = 0; i < G.nr; i++) r2[i] = G.r[i] * G.r[i];
    (void)hipFree(d.r);
    RET(upload(&d.r, r2));
  }
  auto tw = [&](int n, int sign, double2 **dst, int count) -> int {
    std::vector<double> re(n), im(n);
    twiddles(n, sign, re.data(), im.data());
    std::vector<double2> v(count);
    for (int j = 0; j < count; j++) v[j] = make_double2(re[j % n], im[j % n]);
    return upload(dst, v);
  };
  RET(tw(G.nx, -1, &d.tw_z, G.nx));
  RET(tw(G.nth, -1, &d.tw_t, G.nth));
  {
    std::vector<double> re(2 * G.nth), im(2 * G.nth);
    twiddles(2 * G.nth, -1, re.data(), im.data());
    std::vector<double2> v(G.nth + 1);
    for (int j = 0; j <= G.nth; j++) v[j] = make_double2(re[j], im[j]);
    RET(upload(&d.tw_h, v));
    std::vector<double2> v2(2 * (size_t)G.nth);
    for (int j = 0; j < 2 * G.nth; j++) v2[j] = make_double2(re[j], im[j]);
    RET(upload(&d.tw_h2, v2));
  }
  if (!G.hv_r.empty()) {
    std::vector<double> hv;
    hv.insert(hv.end(), G.hv_r.begin(), G.hv_r.end());
    hv.insert(hv.end(), G.hv_t.begin(), G.hv_t.end());
    hv.insert(hv.end(), G.hv_x.begin(), G.hv_x.end());
    hv.insert(hv.end(), G.hd_r.begin(), G.hd_r.end());
    hv.insert(hv.end(), G.hd_t.begin(), G.hd_t.end());
    hv.insert(hv.end(), G.hd_x.begin(), G.hd_x.end());
    RET(upload(&d.hv, hv));
  }
  MLG_CHECK_HIP(hipMalloc((void **)&d.red, sizeof(double) * (4 * (size_t)G.ntchop * G.nxchopdim + 2)));
  MLG_CHECK_HIP(hipMalloc((void **)&d.vproj, sizeof(double) * ((size_t)G.nrchopdim + 1)));
  MLG_CHECK_HIP(hipMalloc((void **)&d.err, sizeof(int)));
  MLG_CHECK_HIP(hipMemset(d.err, 0, sizeof(int)));
  MLG_CHECK_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  *out = c;
  return 0;
}

static void free_field_storage(mlg_field *f) {
  if (f->e) (void)hipFree(f->e);
  if (f->ln) (void)hipFree(f->ln);
  f->e = nullptr;
  f->ln = nullptr;
}

extern "C" int mlg_finalize(mlg_ctx *c) {
  if (!c) return MLG_E_ARG;
  (void)hipStreamSynchronize(c->stream);
  for (mlg_field *f : c->pool) {
    free_field_storage(f);
    delete f;
  }
  c->pool.clear();
  for (mlg_field *f : c->owned) {
    free_field_storage(f);
    delete f;
  }
  c->owned.clear();
  timer_clear(c);
  for (hipEvent_t e : c->timer.spare) (void)hipEventDestroy(e);
  c->timer.spare.clear();
  DevTables &d = c->d;
  void *ptrs[] = {d.pf, d.vdt, d.w, d.ones, d.r, d.x, d.lnx, d.at1, d.ak, d.norm, d.lognorm, d.bxm, d.bxxdx, d.bdel2h,
                  d.tw_z, d.tw_t, d.tw_h, d.tw_h2, d.hv, d.red, d.vproj, d.cvdt, d.err,
                  d.pf_sb, d.pf_eb, d.pf_sf, d.pf_ef, d.vdt_s, d.vdt_e, d.cvdt_s, d.cvdt_e};
  for (auto &f : c->fcache) (void)hipFree(f.f);
  c->fcache.clear();
  for (auto &v : c->m0vecs) (void)hipFree(v.second);
  c->m0vecs.clear();
  for (auto &v : c->adjv) (void)hipFree(v.second);
  c->adjv.clear();
  if (c->d_sponge) (void)hipFree(c->d_sponge);
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return 0;
}

extern "C" int mlg_set_legendre_precision(mlg_ctx *c, int mode) {
  if (!c || (mode != MLG_LEG_FP64 && mode != MLG_LEG_SPLIT32)) return MLG_E_ARG;
  if (mode == MLG_LEG_SPLIT32 && !c->d.pf_sb) RET(build_split_tables(c));
  c->leg_split = mode == MLG_LEG_SPLIT32;
  return 0;
}

extern "C" int mlg_legendre_precision(const mlg_ctx *c) { return c ? (c->leg_split ? MLG_LEG_SPLIT32 : MLG_LEG_FP64) : -1; }

extern "C" int mlg_set_split_fp64_modes(mlg_ctx *c, int nmodes) {
  if (!c || nmodes < 2) return MLG_E_ARG;
  c->split_fp64_modes = nmodes;
  return 0;
}

extern "C" int mlg_dims(const mlg_ctx *c, int *o) {
  if (!c || !o) return MLG_E_ARG;
  const Grid &G = c->g;
  int v[12] = {G.ndimr, G.ndimth, G.ndimx, G.nrchopdim, G.ntchopdim, G.nxchopdim, G.nrh, G.nxchoph, G.nrchop, G.npf, G.nnorm, G.nr};
  std::memcpy(o, v, sizeof(v));
  return 0;
}

extern "C" int mlg_get_table(const mlg_ctx *c, const char *name, double *out, long n, long *len) {
  if (!c || !name) return MLG_E_ARG;
  const Grid &G = c->g;
  const std::vector<double> *v = nullptr;
  std::string s(name);
  if (s == "x") v = &G.x;
  else if (s == "w") v = &G.w;
  else if (s == "r") v = &G.r;
  else if (s == "lnx") v = &G.lnx;
  else if (s == "pf") v = &G.pf;
  else if (s == "lognorm") v = &G.lognorm;
  else if (s == "norm") v = &G.norm;
  else if (s == "at0") v = &G.at0;
  else if (s == "at1") v = &G.at1;
  else if (s == "ak") v = &G.ak;
  else if (s == "vdt") v = &G.vdt;
  else return MLG_E_ARG;
  if (len) *len = (long)v->size();
  if (out) {
    if (n < (long)v->size()) return MLG_E_ARG;
    std::memcpy(out, v->data(), sizeof(double) * v->size());
  }
  return 0;
}

extern "C" int mlg_set_stream(mlg_ctx *c, void *s) {
  if (!c) return MLG_E_ARG;
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return 0;
}

extern "C" int mlg_synchronize(mlg_ctx *c) {
  if (!c) return MLG_E_ARG;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  int e = 0;
  MLG_CHECK_HIP(hipMemcpy(&e, c->d.err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) MLG_CHECK_HIP(hipMemset(c->d.err, 0, sizeof(int)));
  return e;
}

extern "C" int mlg_chopset(mlg_ctx *c, int iof) {
  if (!c) return MLG_E_ARG;
  int v = c->g.nrchop + iof;
  if (v > c->g.nrchopdim || v < 0) return 1;  // CHOPSET: NRCHOP TOO LARGE (mod_scalar3.f90:183-190)
  c->g.nrchop = v;
  return 0;
}
extern "C" int mlg_nrchop(const mlg_ctx *c) { return c ? c->g.nrchop : -1; }

// ------------------------------------------------------------------ fields
static int new_field(mlg_ctx *c, int space, mlg_field **out) {
  mlg_field *f = new mlg_field();
  f->ctx = c;
  f->space = space;
  f->pooled = false;
  if (hipMalloc((void **)&f->e, sizeof(double2) * (size_t)c->n_max) != hipSuccess) {
    delete f;
    return MLG_E_NOMEM;
  }
  if (hipMalloc((void **)&f->ln, sizeof(double) * 2) != hipSuccess) {
    (void)hipFree(f->e);
    delete f;
    return MLG_E_NOMEM;
  }
  // cleared with kernels (no memset nodes if this ever runs under capture)
  RET(launch_zero(c, f->e, c->n_max));
  RET(launch_set_ln(c, f->ln, nullptr, 0.0, 0.0));
  *out = f;
  return 0;
}

extern "C" int mlg_field_create(mlg_ctx *c, int space, mlg_field **f) {
  if (!c || !f || space < 0 || space > 3) return MLG_E_ARG;
  RET(new_field(c, space, f));
  c->owned.insert(*f);
  return 0;
}
extern "C" int mlg_field_destroy(mlg_field *f) {
  if (!f) return MLG_E_ARG;
  mlg_ctx *c = f->ctx;
  if (!c->owned.erase(f)) return MLG_E_ARG;
  (void)hipStreamSynchronize(c->stream);
  free_field_storage(f);
  delete f;
  return 0;
}
extern "C" int mlg_field_space(const mlg_field *f) { return f ? f->space : -1; }

// pooled temporaries (ALLOCATE of a work SCALAR); LN starts at 0
namespace mlg {
mlg_field *pool_take(mlg_ctx *c, int space) {
  mlg_field *f = nullptr;
  if (!c->pool.empty()) {
    f = c->pool.back();
    c->pool.pop_back();
  } else if (new_field(c, space, &f)) {
    return nullptr;
  }
  f->space = space;
  f->pooled = true;
  return f;
}
void pool_give(mlg_ctx *c, mlg_field *f) {
  if (f) c->pool.push_back(f);
}
}  // namespace mlg

static void swap_storage(mlg_field *a, mlg_field *b) { std::swap(a->e, b->e); }

extern "C" int mlg_field_upload(mlg_field *f, const double *e, int space, double ln) {
  if (!f || !e || space < 0 || space > 3) return MLG_E_ARG;
  mlg_ctx *c = f->ctx;
  long n = space_count(c, space);
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(f->e, e, sizeof(double2) * (size_t)n, hipMemcpyHostToDevice));
  MLG_CHECK_HIP(hipMemcpy(f->ln, &ln, sizeof(double), hipMemcpyHostToDevice));
  f->space = space;
  return 0;
}

extern "C" int mlg_field_download(const mlg_field *f, double *e, int *space, double *ln) {
  if (!f) return MLG_E_ARG;
  mlg_ctx *c = f->ctx;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  long n = space_count(c, f->space);
  if (e) {
    MLG_CHECK_HIP(hipMemcpy(e, f->e, sizeof(double2) * (size_t)n, hipMemcpyDeviceToHost));
    if (f->space == MLG_PFP) {  // slots m >= NTCHOPDIM of a PFP field are zero
      const Grid &G = c->g;
      double2 *h = (double2 *)e;
      for (int z = 0; z < G.ndimx; z++)
        for (int j = G.ntchopdim; j < G.ndimth; j++)
          for (int r = 0; r < G.ndimr; r++) h[(size_t)r + (size_t)G.ndimr * ((size_t)j + (size_t)G.ndimth * z)] = make_double2(0, 0);
    }
  }
  if (space) *space = f->space;
  if (ln) MLG_CHECK_HIP(hipMemcpy(ln, f->ln, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mlg_field_device_ptr(const mlg_field *f, void **e, void **ln) {
  if (!f) return MLG_E_ARG;
  if (e) *e = (void *)f->e;
  if (ln) *ln = (void *)f->ln;
  return 0;
}

extern "C" int mlg_field_zero(mlg_field *f, int space) {
  if (!f || space < 0 || space > 3) return MLG_E_ARG;
  f->space = space;
  RET(launch_zero(f->ctx, f->e, f->ctx->n_max));
  return launch_set_ln(f->ctx, f->ln, nullptr, 0.0, 0.0);
}

// COPY0 (mod_fft.f90:706-769): copy, LN included; FFF copies are chopped.
int mlg::copy_field(mlg_field *b, const mlg_field *a) {
  mlg_ctx *c = a->ctx;
  if (a == b) return 0;
  if (a->space == MLG_FFF) {
    b->space = MLG_FFF;
    return launch_copy0_fff(c, b->e, b->ln, a->e, a->ln);
  }
  RET(launch_copy(c, b->e, a->e, space_count(c, a->space)));
  RET(launch_set_ln(c, b->ln, a->ln, 1.0, 0.0));
  b->space = a->space;
  if (a->space == MLG_FFF) RET(launch_chopdo(c, b->e, MLG_FFF));
  return 0;
}
static int copy0(mlg_field *b, const mlg_field *a) { return copy_field(b, a); }
extern "C" int mlg_copy(mlg_field *b, const mlg_field *a) {
  if (!a || !b) return MLG_E_ARG;
  return copy0(b, a);
}
extern "C" int mlg_chopdo(mlg_field *a) {
  if (!a) return MLG_E_ARG;
  return launch_chopdo(a->ctx, a->e, a->space);
}

// ------------------------------------------------------------------ transforms
extern "C" int mlg_rtran(mlg_field *a, int is) {
  if (!a) return MLG_E_ARG;
  mlg_ctx *c = a->ctx;
  Temps T(c);
  mlg_field *t = T.buffer();
  RET(T.ready());
  if (is == -1) {
    if (a->space != MLG_PFF) return 2;
    RET(launch_rtran_fwd(c, a->e, t->e, a->ln));  // writes the truncated rows as zeros, and LN = 0
    swap_storage(a, t);
    a->space = MLG_FFF;
    return 0;
  } else {
    if (a->space != MLG_FFF) return 2;
    RET(launch_rtran_bwd(c, a->e, a->ln, t->e));  // writes the padding rows too
    swap_storage(a, t);
    a->space = MLG_PFF;
  }
  return launch_set_ln(c, a->ln, nullptr, 0.0, 0.0);
}

extern "C" int mlg_verfft(mlg_field *a, int is) {
  if (!a) return MLG_E_ARG;
  mlg_ctx *c = a->ctx;
  Temps T(c);
  mlg_field *t = T.buffer();
  RET(T.ready());
  if (is == -1) {
    if (a->space != MLG_PFP) return 2;
    RET(launch_verfft_fwd(c, a->e, t->e));
    swap_storage(a, t);
    a->space = MLG_PFF;
  } else {
    if (a->space != MLG_PFF) return 2;
    RET(launch_verfft_bwd(c, a->e, t->e));
    swap_storage(a, t);
    a->space = MLG_PFP;
  }
  return 0;
}

extern "C" int mlg_horfft(mlg_field *a, int is) {
  if (!a) return MLG_E_ARG;
  mlg_ctx *c = a->ctx;
  if (is == -1) {
    if (a->space != MLG_PPP) return 2;
    RET(launch_horfft_fwd(c, a->e));
    a->space = MLG_PFP;
  } else {
    if (a->space != MLG_PFP) return 2;
    RET(launch_horfft_bwd(c, a->e));
    a->space = MLG_PPP;
  }
  return 0;
}

extern "C" int mlg_toff(mlg_field *a) {
  if (!a) return MLG_E_ARG;
  if (a->space == MLG_PPP) RET(mlg_horfft(a, -1));
  if (a->space == MLG_PFP) RET(mlg_verfft(a, -1));
  if (a->space == MLG_PFF) RET(mlg_rtran(a, -1));
  return 0;
}
extern "C" int mlg_tofp(mlg_field *a) {
  if (!a) return MLG_E_ARG;
  if (a->space == MLG_FFF) RET(mlg_rtran(a, 1));
  if (a->space == MLG_PFF) RET(mlg_verfft(a, 1));
  if (a->space == MLG_PFP) RET(mlg_horfft(a, 1));
  return 0;
}
extern "C" int mlg_plnfft(mlg_field *a, int is) {
  if (!a) return MLG_E_ARG;
  if (is == -1) {
    if (a->space != MLG_PPP) return 2;
    RET(mlg_horfft(a, -1));
    return mlg_verfft(a, -1);
  }
  if (a->space != MLG_PFF) return 2;
  RET(mlg_verfft(a, 1));
  return mlg_horfft(a, 1);
}

// ------------------------------------------------------------------ operators
// in-place calls (input == output) go through a temporary copy
#define ALIAS_GUARD(inp, outp)                        \
  Temps _T((inp)->ctx);                               \
  const mlg_field *inp##_src = (inp);                 \
  if ((const void *)(inp) == (const void *)(outp)) {  \
    mlg_field *_t = _T.buffer();                      \
    RET(_T.ready());                                  \
    RET(copy0(_t, (inp)));                            \
    inp##_src = _t;                                   \
  }

extern "C" int mlg_delsqh(const mlg_field *a, mlg_field *b) {
  if (!a || !b) return MLG_E_ARG;
  if (a->space != MLG_FFF) return 3;
  ALIAS_GUARD(a, b);
  b->space = MLG_FFF;
  return launch_delsqh(a->ctx, a_src->e, a_src->ln, b->e, b->ln);
}

extern "C" int mlg_idelsqh(const mlg_field *b, mlg_field *a) {
  if (!a || !b) return MLG_E_ARG;
  if (b->space != MLG_FFF) return 3;
  ALIAS_GUARD(b, a);
  a->space = MLG_FFF;
  RET(launch_idelsqh(a->ctx, b_src->e, a->e, a->ln));
  return launch_fix_at_infinity(a->ctx, a->e);
}

extern "C" int mlg_mulxm(const mlg_field *a, mlg_field *b) {
  if (!a || !b) return MLG_E_ARG;
  if (a->space != MLG_FFF) return 3;
  ALIAS_GUARD(a, b);
  b->space = MLG_FFF;
  return launch_band3(a->ctx, 0, a_src->e, a_src->ln, b->e, b->ln);
}

extern "C" int mlg_xxdx(const mlg_field *a, mlg_field *b) {
  if (!a || !b) return MLG_E_ARG;
  if (a->space != MLG_FFF) return 3;
  ALIAS_GUARD(a, b);
  b->space = MLG_FFF;
  return launch_band3(a->ctx, 1, a_src->e, a_src->ln, b->e, b->ln);
}

extern "C" int mlg_del2(const mlg_field *a, mlg_field *b) {
  if (!a || !b) return MLG_E_ARG;
  if (a->space != MLG_FFF) return 3;
  ALIAS_GUARD(a, b);
  b->space = MLG_FFF;
  RET(launch_check_top_m1(a->ctx, a_src->e));
  return launch_del2(a->ctx, a_src->e, a_src->ln, b->e, b->ln);
}

// chop false: A is left without its CHOPDO and fix-up at infinity (for a
// caller whose only reader of A applies them)
static int solve_op(const mlg_field *b, mlg_field *a, int op, double alp, int has_ln, double ln, bool chop = true) {
  if (!a || !b) return MLG_E_ARG;
  if (b->space != MLG_FFF) return 3;
  mlg_ctx *c = b->ctx;
  // A = B then the substitution in A (the reference's COPY + in-place solve):
  // the solve reads B and writes A directly, the CHOPDO below clears the rest
  Temps T(c);
  mlg_field *t = T.buffer();
  RET(T.ready());
  a->space = MLG_FFF;
  // IDEL2/IDEL2LN set A%LN in the solve (the thread of the (m=0, k=0)
  // system); only a truncation without that system needs the zero here
  if (op != 2 && (c->g.nrchop < 1 || c->g.ntchop < 1)) RET(launch_set_ln(c, a->ln, nullptr, 0.0, 0.0));
  RET(launch_band_solve(c, op, b->e, a->e, a->ln, b->ln, alp, has_ln, ln, t->e));
  if (chop) RET(launch_chopdo_fff_n(c, 1, &a->e, op != 2));  // CHOPDO, and the fix at infinity of IDEL2, IDEL2LN
  return 0;
}
extern "C" int mlg_idel2(const mlg_field *b, mlg_field *a, int has_ln, double ln) { return solve_op(b, a, 0, 0.0, has_ln, ln); }
extern "C" int mlg_idel2ln(const mlg_field *b, mlg_field *a) { return solve_op(b, a, 1, 0.0, 0, 0.0); }
extern "C" int mlg_ihelm(const mlg_field *b, mlg_field *a, double alp) {
  if (alp == 0.0) return 3;  // IHELM: ALP CAN'T BE ZERO (mod_legops.f90:1243-1246)
  return solve_op(b, a, 2, alp, 0, 0.0);
}

extern "C" int mlg_ihelmp(const mlg_field *b, mlg_field *a, int p, double alp, double nu) {
  if (!b || !a) return MLG_E_ARG;
  if (b->space != MLG_FFF) return 3;
  mlg_ctx *c = b->ctx;
  a->space = MLG_FFF;
  return launch_ihelmp(c, p, b->e, b->ln, a->e, a->ln, alp, nu);
}

extern "C" int mlg_grad(const mlg_field *a, mlg_field *rr, mlg_field *rp, mlg_field *z) {
  if (!a || !rr || !rp || !z) return MLG_E_ARG;
  if (a->space != MLG_FFF) return 3;
  Temps T(a->ctx);
  const mlg_field *as = a;
  if (a == rr || a == rp || a == z) {
    mlg_field *t = T.buffer();
    RET(T.ready());
    RET(copy0(t, a));
    as = t;
  }
  rr->space = rp->space = z->space = MLG_FFF;
  return launch_grad(a->ctx, as->e, as->ln, rr->e, rr->ln, rp->e, rp->ln, z->e, z->ln);
}

// fix_nrchop >= 1: psi and chi as PROJECT and IDEL2 left them without
// IDEL2's CHOPDO and the fix-ups at infinity, which are applied to the
// values read (launch_pc2vel)
static int pc2vel_impl(const mlg_field *psi, const mlg_field *chi, mlg_field *rur, mlg_field *rup, mlg_field *uz,
                       int fix_nrchop) {
  if (!psi || !chi || !rur || !rup || !uz) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return 3;
  if (rur == psi || rur == chi || rup == psi || rup == chi || uz == psi || uz == chi) return MLG_E_ARG;  // outputs must not alias inputs
  mlg_ctx *c = psi->ctx;
  rur->space = rup->space = uz->space = MLG_FFF;
  return launch_pc2vel(c, psi->e, psi->ln, chi->e, chi->ln, rur->e, rup->e, uz->e, rur->ln, rup->ln, uz->ln, fix_nrchop);
}
extern "C" int mlg_pc2vel(const mlg_field *psi, const mlg_field *chi, mlg_field *rur, mlg_field *rup, mlg_field *uz) {
  return pc2vel_impl(psi, chi, rur, rup, uz, 0);
}

extern "C" int mlg_pc2vor(const mlg_field *psi, const mlg_field *chi, mlg_field *ror, mlg_field *rop, mlg_field *oz) {
  if (!psi || !chi || !ror || !rop || !oz) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return 3;
  if (ror == psi || ror == chi || rop == psi || rop == chi || oz == psi || oz == chi) return MLG_E_ARG;  // outputs must not alias inputs
  mlg_ctx *c = psi->ctx;
  RET(launch_check_top_m1(c, chi->e));
  ror->space = rop->space = oz->space = MLG_FFF;
  return launch_pc2vor(c, psi->e, psi->ln, chi->e, chi->ln, ror->e, rop->e, oz->e, ror->ln, rop->ln, oz->ln);
}

extern "C" int mlg_vprodsub(mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *ror, mlg_field *rop, mlg_field *oz) {
  if (!rur || !rup || !uz || !ror || !rop || !oz) return MLG_E_ARG;
  if (rur->space != MLG_PPP || ror->space != MLG_PPP) return MLG_E_SPACE;
  return launch_vprod(rur->ctx, rur->e, rup->e, uz->e, ror->e, rop->e, oz->e);
}

extern "C" int mlg_vdotpsub(const mlg_field *rar, const mlg_field *rap, const mlg_field *az, const mlg_field *rbr,
                            const mlg_field *rbp, const mlg_field *bz, mlg_field *out) {
  if (!rar || !rap || !az || !rbr || !rbp || !bz || !out) return MLG_E_ARG;
  if (rar->space != MLG_PPP || rbr->space != MLG_PPP) return MLG_E_SPACE;
  mlg_ctx *c = rar->ctx;
  out->space = MLG_PPP;
  RET(launch_zero(c, out->e, c->n_ppp));
  RET(launch_set_ln(c, out->ln, nullptr, 0.0, 0.0));
  return launch_vdot(c, rar->e, rap->e, az->e, rbr->e, rbp->e, bz->e, out->e, 1.0);
}

extern "C" int mlg_vel2pln(mlg_field *rup, mlg_field *psi, int has_ln, double ln) {
  if (!rup || !psi) return MLG_E_ARG;
  if (rup->space != MLG_PFF) return 3;
  psi->space = MLG_FFF;
  return launch_vel2pln(rup->ctx, rup->e, psi->ln, has_ln ? 1 : 0, ln);
}

extern "C" int mlg_vel2pln_noalloc(mlg_field *rup, mlg_field *psi, int has_ln, double ln) {
  if (!rup || !psi) return MLG_E_ARG;
  if (rup->space != MLG_PFF) return 3;
  if (rup->ctx->g.nrchop > VEL2PLN_SCRATCH_MAX) return MLG_E_SIZE;
  psi->space = MLG_FFF;
  return launch_vel2pln_noalloc(rup->ctx, rup->e, psi->ln, has_ln ? 1 : 0, ln);
}

// PROJECT (src/mod_legops.f90:2397-2648).  ln_mode 0: no LN argument; 1: LN
// given as a value; 2: LN given as psi's own LN (the ETD steps pass PSI%LN),
// so the device copy is kept without a host read.
// n fields of one space, FFF or PFP, to PFF with one transform launch (each
// field as mlg_rtran(., 1) or mlg_verfft(., -1) treats it)
// keep[i] (optional, FFF input only) receives the FFF elements of fs[i]
// (not its LN) instead of discarding them; it must be a temporary.
static int to_pff_n(mlg_ctx *c, int n, mlg_field *const *fs, mlg_field *const *keep = nullptr) {
  const int sp = fs[0]->space;
  if (n > MAX_TRANSFORM_JOBS || (sp != MLG_FFF && sp != MLG_PFP)) return 2;
  if (keep && sp != MLG_FFF) return 2;
  for (int i = 1; i < n; i++)
    if (fs[i]->space != sp) return 2;
  Temps T(c);
  mlg_field *t[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) t[i] = T.buffer();
  RET(T.ready());
  const double2 *in[MAX_TRANSFORM_JOBS];
  const double *ln[MAX_TRANSFORM_JOBS];
  double2 *out[MAX_TRANSFORM_JOBS];
  double *lns[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) {
    in[i] = fs[i]->e;
    ln[i] = fs[i]->ln;
    out[i] = t[i]->e;
    lns[i] = fs[i]->ln;
  }
  if (sp == MLG_FFF) {
    RET(launch_rtran_bwd_n(c, n, in, ln, out));  // padding rows included
  } else {
    RET(launch_verfft_fwd_n(c, n, in, out));
  }
  for (int i = 0; i < n; i++) {
    swap_storage(fs[i], t[i]);
    fs[i]->space = MLG_PFF;
    if (keep) {
      swap_storage(keep[i], t[i]);
      keep[i]->space = MLG_FFF;
    }
  }
  if (sp == MLG_FFF) RET(launch_zero_lns(c, lns, n));
  return 0;
}

// keep (optional, FFF velocities only): receives the FFF elements of the
// velocities, see to_pff_n
// fix false (non-FFF velocities only): psi's value-at-infinity fix-up is
// left to the caller
static int project_impl(mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *psi, mlg_field *chi, int ln_mode, double ln,
                        mlg_field *const *keep = nullptr, bool fix = true) {
  if (!(rur->space == rup->space && rup->space == uz->space)) return 2;
  mlg_ctx *c = rur->ctx;
  mlg_field *const vel[3] = {rur, rup, uz};
  if (keep) {
    // spectral velocities that the caller hands over: the composed
    // transform-and-project reads them as they are, and keep[] takes them
    if (rur->space != MLG_FFF) return 2;
    if (psi->space != MLG_FFF) {
      psi->space = MLG_FFF;
      RET(launch_set_ln(c, psi->ln, nullptr, 0.0, 0.0));
    }
    if (chi->space != MLG_FFF) {
      chi->space = MLG_FFF;
      RET(launch_set_ln(c, chi->ln, nullptr, 0.0, 0.0));
    }
    const double *lnu[3] = {rur->ln, rup->ln, uz->ln};
    RET(launch_project_fff(c, rur->e, rup->e, uz->e, lnu, psi->e, psi->ln, chi->e, ln_mode, ln));
    for (int i = 0; i < 3; i++) {
      swap_storage(keep[i], vel[i]);
      std::swap(keep[i]->ln, vel[i]->ln);
      keep[i]->space = MLG_FFF;
    }
    return 0;  // the value-at-infinity fix-up of psi is part of the launch
  } else if (rur->space == MLG_PPP) {
    RET(mlg_plnfft(rur, -1));
    RET(mlg_plnfft(rup, -1));
    RET(mlg_plnfft(uz, -1));
  } else if (rur->space == MLG_PFP || rur->space == MLG_FFF) {
    RET(to_pff_n(c, 3, vel));
  } else if (rur->space != MLG_PFF) {
    return 2;
  }
  if (psi->space != MLG_FFF) {
    psi->space = MLG_FFF;
    RET(launch_set_ln(c, psi->ln, nullptr, 0.0, 0.0));
  }
  if (chi->space != MLG_FFF) {
    chi->space = MLG_FFF;
    RET(launch_set_ln(c, chi->ln, nullptr, 0.0, 0.0));
  }
  if (ln_mode) RET(launch_vel2pln(c, rup->e, psi->ln, ln_mode, ln));
  RET(launch_project(c, rur->e, rup->e, uz->e, psi->e, chi->e));  // CHOPDO of psi, chi included
  return fix ? launch_fix_at_infinity(c, psi->e) : 0;
}

extern "C" int mlg_project(mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *psi, mlg_field *chi, int has_ln, double ln) {
  if (!rur || !rup || !uz || !psi || !chi) return MLG_E_ARG;
  return project_impl(rur, rup, uz, psi, chi, has_ln ? 1 : 0, ln);
}

// ------------------------------------------------------------------ Boussinesq terms
// FFF -> PFF -> PFP (TOFP without the theta transform, done by the fused
// product kernel)
static int to_pfp(mlg_field *a) {
  if (a->space == MLG_FFF) RET(mlg_rtran(a, 1));
  if (a->space == MLG_PFF) RET(mlg_verfft(a, 1));
  return a->space == MLG_PFP ? 0 : 2;
}

// to_pfp of n independent FFF fields: one RTRAN backward launch and one
// VERFFT backward launch for all of them (each field as mlg_rtran then
// mlg_verfft treat it)
// keep[i] (i < nkeep, optional) receives the FFF contents of fs[i] (E and LN)
// instead of discarding them; it must be a temporary of the caller.
static int to_pfp_n(mlg_ctx *c, int n, mlg_field *const *fs, mlg_field *const *keep = nullptr, int nkeep = 0) {
  bool all_fff = n <= MAX_TRANSFORM_JOBS;
  for (int i = 0; i < n; i++) all_fff = all_fff && fs[i]->space == MLG_FFF;
  if (!all_fff) {
    for (int i = 0; i < nkeep; i++) {  // unchopped copies
      if (fs[i]->space != MLG_FFF) return 2;
      RET(launch_copy(c, keep[i]->e, fs[i]->e, space_count(c, MLG_FFF)));
      RET(launch_set_ln(c, keep[i]->ln, fs[i]->ln, 1.0, 0.0));
      keep[i]->space = MLG_FFF;
    }
    for (int i = 0; i < n; i++) RET(to_pfp(fs[i]));
    return 0;
  }
  Temps T(c);
  mlg_field *t[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) t[i] = T.buffer();
  RET(T.ready());
  const double2 *in[MAX_TRANSFORM_JOBS];
  const double *ln[MAX_TRANSFORM_JOBS];
  double2 *out[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) {
    in[i] = fs[i]->e;
    ln[i] = fs[i]->ln;
    out[i] = t[i]->e;
  }
  RET(launch_rtran_bwd_n(c, n, in, ln, out));  // padding rows included
  double *lns[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) {
    swap_storage(fs[i], t[i]);
    if (i < nkeep) {  // the FFF storage and its LN move to keep[i]
      std::swap(keep[i]->e, t[i]->e);
      std::swap(keep[i]->ln, fs[i]->ln);
      keep[i]->space = MLG_FFF;
    }
    fs[i]->space = MLG_PFF;
    lns[i] = fs[i]->ln;
    in[i] = fs[i]->e;
    out[i] = t[i]->e;  // the FFF storage (or keep[i]'s), free again
  }
  RET(launch_verfft_bwd_n(c, n, in, out, 0, 0, lns, n));  // the PFF fields' LN = 0
  for (int i = 0; i < n; i++) {
    swap_storage(fs[i], t[i]);
    fs[i]->space = MLG_PFP;
  }
  return 0;
}

// The seven FFF fields fs to PFP (as to_pfp_n, with keep for the first
// three), together with GRAD's i m b and i k b in PFP (rbp, bz), which come
// from one backward Legendre transform of b (b stays in FFF): the
// z-transform multiplies by i m or i k as it loads.  The same linear maps as
// GRAD then RTRAN, VERFFT, in another order.
static int to_pfp_b(mlg_ctx *c, mlg_field *const *fs, const mlg_field *b, mlg_field *rbp, mlg_field *bz,
                    mlg_field *const *keep) {
  constexpr int n = 7;
  for (int i = 0; i < n; i++)
    if (fs[i]->space != MLG_FFF) return 2;
  Temps T(c);
  mlg_field *t[n + 1];
  for (int i = 0; i <= n; i++) t[i] = T.buffer();
  RET(T.ready());
  const double2 *in[MAX_TRANSFORM_JOBS];
  const double *ln[MAX_TRANSFORM_JOBS];
  double2 *out[MAX_TRANSFORM_JOBS];
  for (int i = 0; i < n; i++) {
    in[i] = fs[i]->e;
    ln[i] = fs[i]->ln;
    out[i] = t[i]->e;
  }
  in[n] = b->e;
  ln[n] = b->ln;
  out[n] = t[n]->e;
  RET(launch_rtran_bwd_n(c, n + 1, in, ln, out));  // padding rows included
  double *lns[n];
  for (int i = 0; i < n; i++) {
    swap_storage(fs[i], t[i]);
    if (keep && i < 3) {  // the FFF storage and its LN move to keep[i]
      std::swap(keep[i]->e, t[i]->e);
      std::swap(keep[i]->ln, fs[i]->ln);
      keep[i]->space = MLG_FFF;
    }
    fs[i]->space = MLG_PFF;
    lns[i] = fs[i]->ln;
    in[i] = fs[i]->e;
    out[i] = t[i]->e;
  }
  in[n] = t[n]->e;  // b in PFF, transformed twice: times i m and times i k
  out[n] = rbp->e;
  in[n + 1] = t[n]->e;
  out[n + 1] = bz->e;
  RET(launch_verfft_bwd_n(c, n + 2, in, out, 1u << (n + 1), 1u << n, lns, n));  // the PFF fields' LN = 0
  for (int i = 0; i < n; i++) {
    swap_storage(fs[i], t[i]);
    fs[i]->space = MLG_PFP;
  }
  rbp->space = bz->space = MLG_PFP;
  return 0;
}

// CALC_BOUSSI_FORCE (src/mod_boussinesq.f90:317-410)
extern "C" int mlg_force(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, mlg_field *psin, mlg_field *chin,
                         mlg_field *bn) {
  if (!psi || !chi || !b || !psin || !chin || !bn) return MLG_E_ARG;
  mlg_ctx *c = psi->ctx;
  Temps T(c);
  mlg_field *rur = T.get(), *rup = T.get(), *uz = T.get(), *ror = T.get(), *rop = T.get(), *oz = T.get();
  mlg_field *rbr = T.get(), *rbp = T.get(), *bz = T.get(), *w = T.get(), *w2 = T.get();
  RET(T.ready());
  RET(mlg_chopset(c, 3));
  int rc = mlg_grad(b, rbr, rbp, bz);
  if (!rc) rc = mlg_pc2vel(psi, chi, rur, rup, uz);
  if (!rc) rc = mlg_pc2vor(psi, chi, ror, rop, oz);
  mlg_field *fs[9] = {rur, rup, uz, ror, rop, oz, rbr, rbp, bz};
  const bool fused = fused_products_supported(c);
  if (!rc && fused) rc = to_pfp_n(c, 9, fs);
  for (int i = 0; i < 9 && !rc && !fused; i++) rc = mlg_tofp(fs[i]);
  mlg_chopset(c, -3);
  RET(rc);
  // W = TOFP(B) at NRCHOP (its value does not depend on the products, so it
  // is formed first and fed to the product kernel)
  RET(copy0(w, b));
  if (fused) {
    RET(to_pfp(w));
    const double2 *in[10] = {rur->e, rup->e, uz->e, ror->e, rop->e, oz->e, rbr->e, rbp->e, bz->e, w->e};
    double2 *out[4] = {ror->e, rop->e, oz->e, bn->e};
    RET(launch_fused_products(c, in, out, 2 * c->g.p.omega, std::pow(c->g.p.bv0, 2.0), 1, bn->ln));  // and N_b's LN = 0
    bn->space = ror->space = rop->space = oz->space = MLG_PFP;
  } else {
    RET(mlg_tofp(w));
    bn->space = MLG_PPP;
    RET(launch_force_products(c, rur->e, rup->e, uz->e, ror->e, rop->e, oz->e, rbr->e, rbp->e, bz->e, w->e, bn->e,
                              2 * c->g.p.omega, std::pow(c->g.p.bv0, 2.0), 1));
    RET(launch_set_ln(c, bn->ln, nullptr, 0.0, 0.0));
    RET(launch_chopdo(c, bn->e, MLG_PPP));
    ror->space = rop->space = oz->space = MLG_PPP;
  }
  RET(mlg_project(ror, rop, oz, psin, w2, 1, 0.0));
  RET(mlg_idel2ln(w2, chin));
  return mlg_toff(bn);
}

// CALC_BOUSSI_NONLIN (src/mod_boussinesq.f90:235-314).  The velocity
// arguments end in physical space, as in the reference, unless
// `keep_velocity` (the time step, which overwrites them next) lets them stay
// in PFP where the fused product kernel leaves them.  With `vel_fff` the
// three spectral velocities given on entry are kept there (FFF, with LN):
// the time step uses them in place of a second PC2VEL of the same PSI, CHI.
static int nonlin_impl(const mlg_field *b, mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *bn, mlg_field *ror,
                       mlg_field *rop, mlg_field *oz, bool keep_velocity, mlg_field *const *vel_fff = nullptr);
extern "C" int mlg_nonlin(const mlg_field *b, mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *bn, mlg_field *ror,
                          mlg_field *rop, mlg_field *oz) {
  if (!b || !rur || !rup || !uz || !bn || !ror || !rop || !oz) return MLG_E_ARG;
  return nonlin_impl(b, rur, rup, uz, bn, ror, rop, oz, false);
}
static int nonlin_impl(const mlg_field *b, mlg_field *rur, mlg_field *rup, mlg_field *uz, mlg_field *bn, mlg_field *ror,
                       mlg_field *rop, mlg_field *oz, bool keep_velocity, mlg_field *const *vel_fff) {
  if (b->space != MLG_FFF) return 3;
  mlg_ctx *c = b->ctx;
  // LN coefficients: GRAD writes those of rbr (rbp, bz: through the fused
  // path only their E is used), VEL2PLN psi's, IDEL2 chi's; W's (and chi's,
  // before IDEL2) are cleared by the first launch
  Temps T(c);
  mlg_field *rbr = T.buffer(), *rbp = T.buffer(), *bz = T.buffer(), *psi = T.buffer(), *chi = T.buffer(), *w = T.buffer();
  RET(T.ready());
  int rc;
  if (fused_products_supported(c) && c->g.nx <= 256 && pow2(c->g.nx)) {
    // GRAD's r d/dr b is formed in FFF; its i m b and i k b components come
    // from b's own transform (to_pfp_b)
    RET(mlg_chopset(c, 1));
    rc = launch_grad(c, b->e, b->ln, rbr->e, rbr->ln, nullptr, nullptr, nullptr, nullptr, w->ln, chi->ln);
    mlg_chopset(c, -1);
    RET(rc);
    rbr->space = MLG_FFF;
    mlg_field *fs7[7] = {rur, rup, uz, ror, rop, oz, rbr};
    RET(to_pfp_b(c, fs7, b, rbp, bz, vel_fff));
    const double2 *in[10] = {rur->e, rup->e, uz->e, ror->e, rop->e, oz->e, rbr->e, rbp->e, bz->e, nullptr};
    double2 *out[4] = {ror->e, rop->e, oz->e, bn->e};
    RET(launch_fused_products(c, in, out, 0.0, 0.0, 0, bn->ln));  // and N_b's LN = 0
    bn->space = ror->space = rop->space = oz->space = MLG_PFP;
    if (!keep_velocity) {
      RET(mlg_horfft(rur, 1));
      RET(mlg_horfft(rup, 1));
      RET(mlg_horfft(uz, 1));
    }
    {  // the z-transforms of PROJECT's inputs and of N_b (TOFF's first part) in one launch
      mlg_field *const pf4[4] = {ror, rop, oz, bn};
      RET(to_pff_n(c, 4, pf4));
    }
    // psi and chi are read by PC2VEL only: PROJECT's fix-up at infinity of
    // psi, and IDEL2's CHOPDO and fix-up of chi, are made there, on the
    // columns it stages
    const int nrchop0 = c->g.nrchop;
    RET(project_impl(ror, rop, oz, psi, w, 1, 0.0, nullptr, false));
    RET(solve_op(w, chi, 0, 0.0, 0, 0.0, false));  // IDEL2
    RET(mlg_chopset(c, 3));
    rc = pc2vel_impl(psi, chi, ror, rop, oz, nrchop0);
    mlg_chopset(c, -3);
    RET(rc);
    return mlg_toff(bn);
  }
  {
    double *lz[3] = {psi->ln, chi->ln, w->ln};
    RET(launch_zero_lns(c, lz, 3));
  }
  RET(mlg_chopset(c, 1));
  rc = mlg_grad(b, rbr, rbp, bz);
  mlg_chopset(c, -1);
  RET(rc);
  mlg_field *fs[9] = {rur, rup, uz, ror, rop, oz, rbr, rbp, bz};
  if (fused_products_supported(c)) {
    RET(to_pfp_n(c, 9, fs, vel_fff, vel_fff ? 3 : 0));
    const double2 *in[10] = {rur->e, rup->e, uz->e, ror->e, rop->e, oz->e, rbr->e, rbp->e, bz->e, nullptr};
    double2 *out[4] = {ror->e, rop->e, oz->e, bn->e};
    RET(launch_fused_products(c, in, out, 0.0, 0.0, 0, bn->ln));  // and N_b's LN = 0
    bn->space = ror->space = rop->space = oz->space = MLG_PFP;
    if (!keep_velocity) {
      RET(mlg_horfft(rur, 1));
      RET(mlg_horfft(rup, 1));
      RET(mlg_horfft(uz, 1));
    }
  } else {
    if (vel_fff)  // unchopped copies: the velocities carry NRCHOP+3 rows
      for (int i = 0; i < 3; i++) {
        RET(launch_copy(c, vel_fff[i]->e, fs[i]->e, space_count(c, MLG_FFF)));
        RET(launch_set_ln(c, vel_fff[i]->ln, fs[i]->ln, 1.0, 0.0));
        vel_fff[i]->space = MLG_FFF;
      }
    for (int i = 0; i < 9; i++) RET(mlg_tofp(fs[i]));
    bn->space = MLG_PPP;
    RET(launch_force_products(c, rur->e, rup->e, uz->e, ror->e, rop->e, oz->e, rbr->e, rbp->e, bz->e, nullptr, bn->e, 0.0,
                              0.0, 0));
    RET(launch_set_ln(c, bn->ln, nullptr, 0.0, 0.0));
    RET(launch_chopdo(c, bn->e, MLG_PPP));
  }
  RET(mlg_project(ror, rop, oz, psi, w, 1, 0.0));
  RET(mlg_idel2(w, chi, 0, 0.0));
  RET(mlg_chopset(c, 3));
  rc = mlg_pc2vel(psi, chi, ror, rop, oz);
  mlg_chopset(c, -3);
  RET(rc);
  return mlg_toff(bn);
}

// CALC_BOUSSI_DIAG + CALC_BOUSSI_ETD_OP (mod_boussinesq.f90:42-109, mod_march.f90:906-945)
extern "C" int mlg_etd_op(const mlg_ctx *c, double dt, double *e16, double *nl16) {
  if (!c || !e16 || !nl16) return MLG_E_ARG;
  typedef std::complex<double> cd;
  const cd I(0.0, 1.0);
  double nf = c->g.p.bv0, om = c->g.p.omega;
  if (nf == 0.0) return 5;  // CALC_BOUSSI_DIAG: N_FREQ = 0, CANNOT DIAGONALIZE
  cd jd[4] = {-I * nf, I * nf, -2.0 * I * om, 2.0 * I * om};
  cd s[16], si[16];
  for (int q = 0; q < 16; q++) s[q] = si[q] = 0.0;
#define M4(a, i, j) a[(i) + 4 * (j)]
  M4(s, 0, 2) = I; M4(s, 0, 3) = -I; M4(s, 1, 2) = 1.0; M4(s, 1, 3) = 1.0;
  M4(s, 2, 0) = -I / nf; M4(s, 2, 1) = I / nf; M4(s, 3, 0) = 1.0; M4(s, 3, 1) = 1.0;
  M4(si, 0, 2) = I * nf / 2.0; M4(si, 0, 3) = 0.5; M4(si, 1, 2) = -I * nf / 2.0; M4(si, 1, 3) = 0.5;
  M4(si, 2, 0) = -I / 2.0; M4(si, 2, 1) = 0.5; M4(si, 3, 0) = I / 2.0; M4(si, 3, 1) = 0.5;
  cd ex[4], ph[4];
  for (int i = 0; i < 4; i++) {
    ex[i] = std::exp(jd[i] * dt);
    ph[i] = (std::abs(jd[i]) > 1.0e-14) ? (ex[i] - 1.0) / jd[i] : cd(dt, 0.0);
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      cd ae = 0.0, an = 0.0;
      for (int l = 0; l < 4; l++) {
        ae += M4(s, i, l) * ex[l] * M4(si, l, j);
        an += M4(s, i, l) * ph[l] * M4(si, l, j);
      }
      e16[i + 4 * j] = ae.real();
      nl16[i + 4 * j] = an.real();
    }
#undef M4
  return 0;
}

// CALC_BOUSSI_ENERGY (src/mod_boussinesq.f90:138-232)
extern "C" int mlg_energy(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, double *ke, double *pe, double *dke) {
  if (!psi || !chi || !b || !ke || !pe || !dke) return MLG_E_ARG;
  mlg_ctx *c = psi->ctx;
  const Grid &G = c->g;
  double bln = 0.0;
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(&bln, b->ln, sizeof(double), hipMemcpyDeviceToHost));
  if (bln != 0.0) return 13;
  Temps T(c);
  mlg_field *w1 = T.get(), *w2 = T.get(), *w3 = T.get();
  RET(T.ready());
  const int nt = G.ntchop, nxc = G.nxchop, nxd = G.nxchopdim;
  const size_t na = (size_t)nt * nxd;
  double *dall = c->d.red;  // [4][nt*nxd]
  RET(launch_zero(c, (double2 *)dall, (long)(4 * na / 2 + 1)));
  RET(mlg_chopset(c, 3));
  int rc = copy0(w1, psi);
  if (!rc) rc = launch_set_ln(c, w1->ln, psi->ln, 2.0, 0.0);
  if (!rc) rc = mlg_delsqh(w1, w2);
  if (!rc) rc = launch_set_ln(c, w1->ln, nullptr, 0, 0);
  if (!rc) rc = launch_set_ln(c, w2->ln, nullptr, 0, 0);
  if (!rc) rc = mlg_rtran(w1, 1);
  if (!rc) rc = mlg_rtran(w2, 1);
  if (!rc) rc = launch_energy_terms(c, w1->e, w2->e, 0, dall);
  if (!rc) rc = mlg_del2(chi, w1);
  if (!rc) rc = mlg_delsqh(chi, w2);
  if (!rc) rc = launch_set_ln(c, w1->ln, nullptr, 0, 0);
  if (!rc) rc = launch_set_ln(c, w2->ln, nullptr, 0, 0);
  if (!rc) rc = mlg_rtran(w1, 1);
  if (!rc) rc = mlg_rtran(w2, 1);
  if (!rc) rc = launch_energy_terms(c, w1->e, w2->e, 0, dall + na);
  if (!rc) rc = copy0(w3, b);
  if (!rc) rc = launch_set_ln(c, w3->ln, nullptr, 0, 0);
  if (!rc) rc = mlg_rtran(w3, 1);
  if (!rc) rc = launch_energy_terms(c, w3->e, w2->e, 0, dall + 2 * na);
  if (!rc) rc = launch_energy_terms(c, w3->e, nullptr, 1, dall + 3 * na);
  mlg_chopset(c, -3);
  RET(rc);
  std::vector<double> h(4 * na);
  MLG_CHECK_HIP(hipStreamSynchronize(c->stream));
  MLG_CHECK_HIP(hipMemcpy(h.data(), dall, sizeof(double) * 4 * na, hipMemcpyDeviceToHost));
  // fold +-k into the NXCHOP non-negative wavenumbers (PRODCT_MK_HALFK)
  auto fold = [&](const double *all, double *out) {
    for (int i = 0; i < nt * nxc; i++) out[i] = 0.0;
    for (int mm = 0; mm < nt; mm++) out[mm] = all[mm];
    for (int kk = 2; kk <= nxc; kk++) {
      int kc = nxd + 2 - kk;
      for (int mm = 0; mm < nt; mm++) out[mm + (size_t)nt * (kk - 1)] = all[mm + (size_t)nt * (kc - 1)] + all[mm + (size_t)nt * (kk - 1)];
    }
  };
  std::vector<double> t1(nt * nxc), t2(nt * nxc), t4(nt * nxc);
  fold(h.data(), t1.data());
  fold(h.data() + na, t2.data());
  fold(h.data() + 2 * na, dke);
  fold(h.data() + 3 * na, t4.data());
  for (int i = 0; i < nt * nxc; i++) {
    ke[i] = (-t1[i] + t2[i]) / 2.0;
    pe[i] = t4[i] / 2.0 / std::pow(G.p.bv0, 2.0);
  }
  return mlg_synchronize(c);
}

// ------------------------------------------------------------------ time march
extern "C" int mlg_solver_destroy(mlg_solver *s);
extern "C" int mlg_solver_create(mlg_ctx *c, mlg_solver **out) {
  if (!c || !out) return MLG_E_ARG;
  mlg_solver *s = new mlg_solver();
  s->ctx = c;
  mlg_field **fs[9] = {&s->n_psi, &s->n_chi, &s->n_b, &s->r_n_r, &s->r_n_p, &s->n_z, &s->r_pi_r, &s->r_pi_p, &s->pi_z};
  for (auto f : fs) {
    int rc = new_field(c, MLG_FFF, f);
    if (rc) {
      mlg_solver_destroy(s);
      return rc;
    }
  }
  *out = s;
  return 0;
}
static void drop_graph(mlg_solver *s);
extern "C" int mlg_solver_destroy(mlg_solver *s) {
  if (!s) return MLG_E_ARG;
  drop_graph(s);
  mlg_field *fs[9] = {s->n_psi, s->n_chi, s->n_b, s->r_n_r, s->r_n_p, s->n_z, s->r_pi_r, s->r_pi_p, s->pi_z};
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto f : fs)
    if (f) {
      free_field_storage(f);
      delete f;
    }
  delete s;
  return 0;
}
extern "C" double mlg_solver_time(const mlg_solver *s) { return s ? s->t : 0.0; }
extern "C" int mlg_solver_nstep(const mlg_solver *s) { return s ? s->n : -1; }

// y.e = x1 + a x2 (+ y.ln likewise)
static int euler(mlg_field *y, const mlg_field *x1, double a, const mlg_field *x2, bool ln) {
  mlg_ctx *c = y->ctx;
  RET(launch_axpby(c, 0, y->e, a, x1->e, x2->e, nullptr, c->n_fff));
  if (ln) RET(launch_ln_axpby(c, 0, y->ln, a, x1->ln, x2->ln, nullptr));
  y->space = MLG_FFF;
  return 0;
}
// y = x1 + dt (1.5 x2 - 0.5 x3)
static int ab2(mlg_field *y, double dt, const mlg_field *x2, const mlg_field *x3, bool ln) {
  mlg_ctx *c = y->ctx;
  RET(launch_axpby(c, 1, y->e, dt, y->e, x2->e, x3->e, c->n_fff));
  if (ln) RET(launch_ln_axpby(c, 1, y->ln, dt, y->ln, x2->ln, x3->ln));
  return 0;
}
// y = 2 x1 - y
static int richardson(mlg_field *y, const mlg_field *x1, bool ln) {
  mlg_ctx *c = y->ctx;
  RET(launch_axpby(c, 2, y->e, 0.0, x1->e, nullptr, nullptr, c->n_fff));
  if (ln) RET(launch_ln_axpby(c, 2, y->ln, 0.0, x1->ln, nullptr, nullptr));
  return 0;
}

// IHELM of n independent fields, A_i = IHELM(B_i, ALP_i), with one band-solve
// launch for all of them (each field as solve_op does it for op 2).
// bil: the right-hand sides b are row-interleaved (launch_visc2_rhs_il)
static int ihelm_n(int n, mlg_field *const *b, mlg_field *const *a, const double *alp, int bil = 0) {
  mlg_ctx *c = b[0]->ctx;
  const double2 *be[3];
  double2 *ae[3];
  double *aln[3];
  const double *bln[3];
  for (int i = 0; i < n; i++) {
    if (alp[i] == 0.0) return 3;  // IHELM: ALP CAN'T BE ZERO (mod_legops.f90:1243-1246)
    if (b[i]->space != MLG_FFF) return 3;
    a[i]->space = MLG_FFF;  // A = B, solved in place: read from B, written to A
    be[i] = b[i]->e;
    ae[i] = a[i]->e;
    aln[i] = a[i]->ln;
    bln[i] = b[i]->ln;
  }
  Temps T(c);
  mlg_field *tf[3] = {nullptr, nullptr, nullptr};
  for (int i = 0; i < n; i++) tf[i] = T.buffer();
  RET(T.ready());
  double2 *te[3] = {nullptr, nullptr, nullptr};
  for (int i = 0; i < n; i++) te[i] = tf[i]->e;
  RET(launch_band_solve_n(c, 2, n, be, ae, aln, bln, alp, 0, 0.0, te, bil));
  return launch_chopdo_fff_n(c, n, ae, false);
}

// VISC1 (mod_march.f90:1529-1576): backward Euler diffusion; the right-hand
// side WK = -ALP*A, the solve A = IHELM(WK, ALP) follows (batched)
static int visc1_rhs(mlg_field *a, double nu, double dt, mlg_field *wk, double *alp) {
  if (nu == 0.0) return 5;
  mlg_ctx *c = a->ctx;
  *alp = 1 / (nu * dt);
  wk->space = MLG_FFF;
  RET(launch_set_ln(c, wk->ln, a->ln, -*alp, 0.0));
  RET(launch_copy(c, wk->e, a->e, c->n_fff));
  return launch_scale(c, wk->e, -*alp, c->n_fff);
}


// psi, chi (NU) and b (KAPPA, when non-zero) are independent: their right-
// hand sides are formed in turn, then the three IHELM solves share a launch.
static int visc_be(mlg_ctx *c, mlg_field *psi, mlg_field *chi, mlg_field *b, double dt) {
  if (c->g.p.visc_sw != 1) return 0;
  Temps T(c);
  mlg_field *wk[3] = {T.get(), T.get(), T.get()};
  RET(T.ready());
  mlg_field *f[3] = {psi, chi, b};
  const double nu[3] = {c->g.p.nu, c->g.p.nu, c->g.p.kappa};
  const int n = c->g.p.kappa != 0.0 ? 3 : 2;
  double alp[3];
  for (int i = 0; i < n; i++) RET(visc1_rhs(f[i], nu[i], dt, wk[i], &alp[i]));
  return ihelm_n(n, wk, f, alp);
}

// VISC2 on (psi, chi, b): the current values are read from fa[] (which may
// be the outputs f[] themselves), the previous ones from fo[]; all three
// right-hand sides are formed before the batched solves write f[].
static int visc_cn(mlg_ctx *c, mlg_field *const *f, const mlg_field *const *fa, const mlg_field *const *fo) {
  if (c->g.p.visc_sw != 1) return 0;
  Temps T(c);
  mlg_field *wk[3] = {T.buffer(), T.buffer(), T.buffer()};  // the right-hand-side kernel writes E and LN
  RET(T.ready());
  const double nu[3] = {c->g.p.nu, c->g.p.nu, c->g.p.kappa};
  const int n = c->g.p.kappa != 0.0 ? 3 : 2;
  double alp[3];
  // VISC2 (mod_march.f90:1579-1636): Crank-Nicolson diffusion, right-hand
  // sides WK = -DEL2(AO) - ALP*A (CHOPSET(2); DEL2; CHOPSET(-2); CHOPDO; axpy
  // in one pass) of all fields in one launch, then A = IHELM(WK, ALP)
  Visc2Jobs jobs;
  for (int i = 0; i < n; i++) {
    if (nu[i] == 0.0) return 5;
    if (fo[i]->space != MLG_FFF) return 3;  // DEL2's input check
    alp[i] = 2 / nu[i] / c->g.p.dt;
    wk[i]->space = MLG_FFF;
    jobs.j[i] = Visc2Job{fo[i]->e, fo[i]->ln, fa[i]->e, fa[i]->ln, alp[i], wk[i]->e, wk[i]->ln};
  }
  // WK row-interleaved for the one-thread-per-system solve, in columns for the partitioned one
  const int il = band_lanes(kgrid(c)) == 0;
  RET(launch_visc2_rhs_il(c, n, jobs, il));
  mlg_field *out[3] = {f[0], f[1], f[2]};
  return ihelm_n(n, wk, out, alp, il);
}
static int visc_cn(mlg_ctx *c, mlg_field *psi, mlg_field *chi, mlg_field *b, const mlg_field *psio, const mlg_field *chio,
                   const mlg_field *bo) {
  mlg_field *const f[3] = {psi, chi, b};
  const mlg_field *const fo[3] = {psio, chio, bo};
  return visc_cn(c, f, f, fo);
}

static bool filter_active(const mlg_ctx *c) { return c->g.p.visc_sw == 3 || c->g.p.hyper_cn; }

// CALC_BOUSSI_HYPERV, SW=3 branch (mod_march.f90:1339-1343)
// HYPERV (src/mod_march.f90:1639-1686): implicit Euler step of dA/dt =
// nu del^2 A - (-1)^(P/2) nup del^P A, A <- IHELMP(P, alp A, -alp, bet)
// with alp = (-1)^(P/2) / (dt nup), bet = -(-1)^(P/2) nu / nup.
static int hyperv2(mlg_ctx *c, mlg_field *a, double nu, double nup, int p, double dt) {
  if (nup == 0.0 || a->space != MLG_FFF) return 5;
  const double sg = ((p / 2) & 1) ? -1.0 : 1.0;
  const double alp = 1 / dt / nup * sg, bet = -sg * nu / nup;
  Temps T(c);
  mlg_field *wk = T.buffer();
  RET(T.ready());
  RET(launch_copy(c, wk->e, a->e, space_count(c, MLG_FFF)));
  RET(launch_scale(c, wk->e, alp, space_count(c, MLG_FFF)));
  RET(launch_set_ln(c, wk->ln, a->ln, alp, 0.0));
  return launch_ihelmp(c, p, wk->e, wk->ln, a->e, a->ln, -alp, bet);
}

// CALC_BOUSSI_HYPERV (src/mod_march.f90:1305-1349): SW = 2, HYPERV of psi,
// chi and (with a hyperdiffusivity) b; SW = 3, the exponential filter.
static int hyperv_state(mlg_ctx *c, mlg_field *psi, mlg_field *chi, mlg_field *b, double dt) {
  if (c->g.p.visc_sw == 2) {
    const mlg_params &P = c->g.p;
    RET(hyperv2(c, psi, P.nu, P.nup, P.visc_p, dt));
    RET(hyperv2(c, chi, P.nu, P.nup, P.visc_p, dt));
    if (P.kappap != 0.0) RET(hyperv2(c, b, P.kappa, P.kappap, P.visc_p, dt));
    return 0;
  }
  if (c->g.p.visc_sw != 3) return 0;
  double tf = dt / c->g.p.dt;
  RET(launch_hyp3(c, psi->e, 0, tf));
  RET(launch_hyp3(c, chi->e, 0, tf));
  return launch_hyp3(c, b->e, 1, tf);
}
// RICH_BOUSSI_FE_BE (mod_march.f90:535-606)
static int rich_fe_be(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  double dt = c->g.p.dt, hdt = dt * 0.5;
  s->t += dt;
  s->n += 1;
  Temps T(c);
  mlg_field *psi2 = T.get(), *chi2 = T.get(), *b2 = T.get(), *pn = T.get(), *cn = T.get(), *bnn = T.get();
  RET(T.ready());
  RET(mlg_force(psi, chi, b, s->n_psi, s->n_chi, s->n_b));
  RET(euler(psi2, psi, hdt, s->n_psi, true));
  RET(euler(chi2, chi, hdt, s->n_chi, true));
  RET(euler(b2, b, hdt, s->n_b, false));
  RET(visc_be(c, psi2, chi2, b2, hdt));
  RET(hyperv_state(c, psi2, chi2, b2, hdt));
  RET(mlg_force(psi2, chi2, b2, pn, cn, bnn));
  RET(euler(psi2, psi2, hdt, pn, true));
  RET(euler(chi2, chi2, hdt, cn, true));
  RET(euler(b2, b2, hdt, bnn, false));
  RET(visc_be(c, psi2, chi2, b2, hdt));
  RET(hyperv_state(c, psi2, chi2, b2, hdt));
  RET(euler(psi, psi, dt, s->n_psi, true));
  RET(euler(chi, chi, dt, s->n_chi, true));
  RET(euler(b, b, dt, s->n_b, false));
  RET(visc_be(c, psi, chi, b, dt));
  RET(hyperv_state(c, psi, chi, b, dt));
  RET(richardson(psi, psi2, true));
  RET(richardson(chi, chi2, true));
  return richardson(b, b2, false);
}

// STEP_BOUSSI_AB_CN (mod_march.f90:609-661)
static int step_ab_cn(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  double dt = c->g.p.dt;
  s->t += dt;
  s->n += 1;
  Temps T(c);
  mlg_field *psin = T.get(), *chin = T.get(), *bn = T.get(), *psi0 = T.get(), *chi0 = T.get(), *b0 = T.get();
  RET(T.ready());
  RET(mlg_force(psi, chi, b, psin, chin, bn));
  // PSI0 = PSI ; AB2 ; N history = the new N, for the three fields in one launch
  mlg_field *y[3] = {psi, chi, b}, *y0[3] = {psi0, chi0, b0}, *h[3] = {s->n_psi, s->n_chi, s->n_b};
  const mlg_field *x[3] = {psin, chin, bn};
  bool fff = true;
  for (int i = 0; i < 3; i++) fff = fff && y[i]->space == MLG_FFF && x[i]->space == MLG_FFF;
  if (fff) {
    Ab2Jobs jobs;
    for (int i = 0; i < 3; i++) {
      y0[i]->space = h[i]->space = MLG_FFF;
      jobs.j[i] = Ab2Job{y[i]->e, y0[i]->e, h[i]->e, x[i]->e, y[i]->ln, y0[i]->ln, h[i]->ln, x[i]->ln, i < 2};
    }
    RET(launch_ab2_hist(c, 3, jobs, dt));
  } else {
    for (int i = 0; i < 3; i++) RET(copy0(y0[i], y[i]));
    for (int i = 0; i < 3; i++) RET(ab2(y[i], dt, x[i], h[i], i < 2));
    for (int i = 0; i < 3; i++) RET(copy0(h[i], x[i]));
  }
  RET(visc_cn(c, psi, chi, b, psi0, chi0, b0));
  return hyperv_state(c, psi, chi, b, dt);
}

// CHOPSET(3); PC2VEL(PSI, CHI, u[0..2]); PC2VOR(PSI, CHI, u[3..5]);
// CHOPSET(-3) in one pass over psi and chi
static int pc2velvor3(mlg_ctx *c, const mlg_field *psi, const mlg_field *chi, mlg_field *const *u) {
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return 3;
  VelVorOut o;
  for (int i = 0; i < 6; i++) {
    if (u[i] == psi || u[i] == chi) return MLG_E_ARG;  // outputs must not alias inputs
    o.e[i] = u[i]->e;
    o.ln[i] = u[i]->ln;
    u[i]->space = MLG_FFF;
  }
  RET(mlg_chopset(c, 3));
  int rc = launch_pc2velvor(c, psi->e, psi->ln, chi->e, chi->ln, o);  // with PC2VOR's DEL2 check
  mlg_chopset(c, -3);
  return rc;
}

// PC2VEL at NRCHOP+3 followed by the pressure-gradient history update
// (R_PI_R = R_PI_R - RUR, ... with hmode 4; R_PI_R = ha (R_PI_R - RUR), ...
// with hmode 3), for velocities used by nothing else: one pass, the
// velocities are not stored
// With `base`, the histories' previous values are read from base (the
// values the ETD update would have stored in them).
static int pc2vel3_hist(mlg_solver *s, const mlg_field *psi, const mlg_field *chi, int hmode, double ha,
                        mlg_field *const *base = nullptr) {
  mlg_ctx *c = s->ctx;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return 3;
  VelTriple bt{{nullptr, nullptr, nullptr}};
  if (base)
    for (int i = 0; i < 3; i++) bt.e[i] = base[i]->e;
  RET(mlg_chopset(c, 3));
  int rc = launch_pc2vel_hist(c, psi->e, psi->ln, chi->e, chi->ln, s->r_pi_r->e, s->r_pi_p->e, s->pi_z->e, hmode, ha, bt);
  mlg_chopset(c, -3);
  return rc;
}

// With the axial sponge on (RMV%SW = 2) the ETD update leaves u and b
// unfiltered: CALC_BOUSSI_ETD_SPONGE(UZN, B) then CALC_BOUSSI_ETD_HYPERV(
// RURN, RUPN, UZN, B), in the reference's order (mod_march.f90:708-709,
// 865-866)
static int sponge_then_filter(mlg_ctx *c, mlg_field *rurn, mlg_field *rupn, mlg_field *uzn, mlg_field *b, double dt) {
  RET(etd_sponge(c, uzn, b, dt));
  if (!filter_active(c)) return 0;
  const double tf = dt / c->g.p.dt;
  RET(launch_hyp3(c, rurn->e, 0, tf));
  RET(launch_hyp3(c, rupn->e, 0, tf));
  RET(launch_hyp3(c, uzn->e, 0, tf));
  return launch_hyp3(c, b->e, 1, tf);
}

// STEP_BOUSSI_ETDFE_BE (mod_march.f90:664-753)
static int step_etdfe_be(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, const double *e16, const double *nl16,
                         double dt, bool save) {
  mlg_ctx *c = s->ctx;
  Temps T(c);
  mlg_field *rurn = T.get(), *rupn = T.get(), *uzn = T.get(), *rorn = T.get(), *ropn = T.get(), *ozn = T.get();
  mlg_field *bn = T.get(), *w = T.get(), *ur = T.buffer(), *up = T.buffer(), *uz = T.buffer();
  mlg_field *const hk[3] = {T.buffer(), T.buffer(), T.buffer()};
  RET(T.ready());
  {
    mlg_field *const u[6] = {rurn, rupn, uzn, rorn, ropn, ozn};
    RET(pc2velvor3(c, psi, chi, u));
  }
  // the second PC2VEL of the reference step equals the first (same PSI, CHI,
  // NRCHOP+3): NONLIN hands the spectral velocities back in ur, up, uz
  mlg_field *const vel[3] = {ur, up, uz};
  RET(nonlin_impl(b, rurn, rupn, uzn, bn, rorn, ropn, ozn, true, vel));
  rurn = ur, rupn = up, uzn = uz;
  // ETD_HYPERV(N) ; ETDFE_SUB ; ETD_HYPERV(u, b) ; history stores when
  // saving: the N history here; the pi history R_PI = u is kept in hk (the
  // velocities' FFF elements, which PROJECT transforms) and combined with the
  // second PC2VEL below
  const bool sponge = c->rmv_sw == 2;
  RET(launch_etd_sub(c, e16, nl16, dt, rurn->e, rupn->e, uzn->e, b->e, rorn->e, ropn->e, ozn->e, bn->e, s->r_pi_r->e,
                     s->r_pi_p->e, s->pi_z->e, s->r_n_r->e, s->r_n_p->e, s->n_z->e, s->n_b->e, 0, save ? 1 : 0, nullptr,
                     nullptr, filter_active(c), dt / c->g.p.dt, nullptr, nullptr, sponge ? 0 : 1));
  if (sponge) RET(sponge_then_filter(c, rurn, rupn, uzn, b, dt));
  RET(project_impl(rurn, rupn, uzn, psi, w, 2, 0.0, hk));  // PROJECT(..., PSI, W, PSI%LN)
  RET(mlg_idel2ln(w, chi));
  if (save) {
    RET(pc2vel3_hist(s, psi, chi, 3, 0.5, hk));
  }
  return 0;
}

// RICH_BOUSSI_ETDFE_BE (mod_march.f90:756-808)
static int rich_etdfe_be(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  double dt = c->g.p.dt, hdt = dt * 0.5;
  double he[16], hn[16];
  RET(mlg_etd_op(c, hdt, he, hn));
  s->t += dt;
  s->n += 1;
  Temps T(c);
  mlg_field *psi2 = T.get(), *chi2 = T.get(), *b2 = T.get();
  RET(T.ready());
  RET(copy0(psi2, psi));
  RET(copy0(chi2, chi));
  RET(copy0(b2, b));
  RET(step_etdfe_be(s, psi2, chi2, b2, he, hn, hdt, false));
  RET(visc_be(c, psi2, chi2, b2, hdt));
  RET(step_etdfe_be(s, psi2, chi2, b2, he, hn, hdt, false));
  RET(visc_be(c, psi2, chi2, b2, hdt));
  RET(step_etdfe_be(s, psi, chi, b, s->etd_e, s->etd_nl, dt, true));
  RET(visc_be(c, psi, chi, b, dt));
  RET(richardson(psi, psi2, true));
  RET(richardson(chi, chi2, true));
  return richardson(b, b2, false);
}

// STEP_BOUSSI_ETDAB_CN (mod_march.f90:811-901)
static int step_etdab_cn(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  double dt = c->g.p.dt;
  s->t += dt;
  s->n += 1;
  // No LN is cleared here: PC2VEL+PC2VOR writes those of rurn .. ozn, NONLIN
  // that of bn, and the ETD launch sets W%LN = 0 before PROJECT
  Temps T(c);
  mlg_field *rurn = T.buffer(), *rupn = T.buffer(), *uzn = T.buffer(), *rorn = T.buffer(), *ropn = T.buffer(),
            *ozn = T.buffer();
  mlg_field *bn = T.buffer(), *w = T.buffer();
  mlg_field *ur = T.buffer(), *up = T.buffer(), *uz = T.buffer(), *psi1 = T.buffer(), *chi1 = T.buffer(), *b1 = T.buffer();
  mlg_field *const hk[3] = {T.buffer(), T.buffer(), T.buffer()};
  RET(T.ready());
  {
    mlg_field *const u[6] = {rurn, rupn, uzn, rorn, ropn, ozn};
    RET(pc2velvor3(c, psi, chi, u));
  }
  // the second PC2VEL of the reference step equals the first (same PSI, CHI,
  // NRCHOP+3): NONLIN hands the spectral velocities back in ur, up, uz
  mlg_field *const vel[3] = {ur, up, uz};
  RET(nonlin_impl(b, rurn, rupn, uzn, bn, rorn, ropn, ozn, true, vel));
  rurn = ur, rupn = up, uzn = uz;
  // B0 = B ; PSI0 = PSI ; CHI0 = CHI ; the ETD update, PROJECT ; VISC2: the
  // new b and the projection are written to b1, psi1, chi1, so that b, psi,
  // chi themselves stay the previous values VISC2 reads, and the solves
  // write the new ones into them.  Without VISC2 they are updated in place.
  const bool cn = c->g.p.visc_sw == 1;
  mlg_field *bj = b, *pj = psi, *cj = chi;
  double *lnd[3] = {nullptr, nullptr, w->ln};  // W%LN = 0
  const double *lns[3] = {nullptr, nullptr, nullptr};
  if (cn) {  // b1, psi1 take the LN of b, psi (copied by the ETD launch)
    bj = b1, pj = psi1, cj = chi1;
    bj->space = pj->space = MLG_FFF;
    lnd[0] = bj->ln, lns[0] = b->ln;
    lnd[1] = pj->ln, lns[1] = psi->ln;
  }
  // ETD_HYPERV(N) ; ETDAB_SUB ; N history ; ETD_HYPERV(u, b) ; pi history:
  // R_PI = u is kept in hk (the velocities' FFF elements, which PROJECT
  // transforms) and combined with the second PC2VEL below
  const bool sponge = c->rmv_sw == 2;
  RET(launch_etd_sub(c, s->etd_e, s->etd_nl, dt, rurn->e, rupn->e, uzn->e, b->e, rorn->e, ropn->e, ozn->e, bn->e, s->r_pi_r->e,
                     s->r_pi_p->e, s->pi_z->e, s->r_n_r->e, s->r_n_p->e, s->n_z->e, s->n_b->e, 1, 1, nullptr, bj->e,
                     filter_active(c), dt / c->g.p.dt, lnd, lns, sponge ? 0 : 1));
  if (sponge) RET(sponge_then_filter(c, rurn, rupn, uzn, bj, dt));
  RET(project_impl(rurn, rupn, uzn, pj, w, 2, 0.0, hk));  // PROJECT(..., PSI, W, PSI%LN)
  RET(mlg_idel2ln(w, cj));
  RET(pc2vel3_hist(s, pj, cj, 4, 0.0, hk));
  if (!cn) return 0;
  mlg_field *const f[3] = {psi, chi, b};
  const mlg_field *const fa[3] = {pj, cj, bj}, *const fo[3] = {psi, chi, b};
  return visc_cn(c, f, fa, fo);
}

extern "C" int mlg_initialize(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  if (!s || !psi || !chi || !b) return MLG_E_ARG;
  mlg_ctx *c = s->ctx;
  // SOLVER_T%INITIALIZE (mod_march.f90:185-201): ETD takes only the
  // exponential hyperviscosity; VISC%SW = 2 aborts with ERR_FLAGS%SOLVER
  if (!c->g.p.adams && c->g.p.visc_sw == 2) return 5;  // mod_march.f90:190-201
  if (c->g.p.visc_sw == 2 && (c->g.p.visc_p < 4 || c->g.p.visc_p > 14 || c->g.p.visc_p % 2)) return 3;
  int rc;
  if (c->g.p.adams) {
    rc = rich_fe_be(s, psi, chi, b);
  } else {
    RET(mlg_etd_op(c, c->g.p.dt, s->etd_e, s->etd_nl));
    rc = rich_etdfe_be(s, psi, chi, b);
  }
  RET(rc);
  return mlg_synchronize(c);
}

static int one_step(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  RET(c->g.p.adams ? step_ab_cn(s, psi, chi, b) : step_etdab_cn(s, psi, chi, b));
  return launch_diagnost(c, psi->e);
}

// DIAGNOST's per-step checks (src/mod_march.f90:1726-1791) on the given
// fields: 11 when psi has coefficients beyond NRCHOPS(m) at m = 0, 1; 12
// when |psi| > 100 in the lowest modes (n < 5, 1 <= m < 6, k < 5).  An error
// raised earlier on the device and not yet read is returned first.
extern "C" int mlg_diagnost(const mlg_field *psi, const mlg_field *chi) {
  if (!psi || !chi) return MLG_E_ARG;
  if (psi->space != MLG_FFF || chi->space != MLG_FFF) return 3;
  mlg_ctx *c = psi->ctx;
  RET(mlg_synchronize(c));
  RET(launch_diagnost(c, psi->e));
  return mlg_synchronize(c);
}

int mlg::step_once(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) { return one_step(s, psi, chi, b); }

// TIME_STEPPING's events after a step (src/mod_march.f90:264-287): REMOVE
// and the Richardson restart every RMV%INT steps, then HYP3ADJ and HYP3SET
// every ADJINT steps (mlg_set_march_controls).  The steps' DIAGNOST codes
// are read first: the reference stops there.
static bool rmv_due(const mlg_solver *s) { return s->rmv_sw != 0 && s->rmv_int > 0 && s->n % s->rmv_int == 0; }
static bool adj_due(const mlg_solver *s) { return s->adj_sw != 0 && s->adj_int > 0 && s->n % s->adj_int == 0; }
static int march_events(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  mlg_ctx *c = s->ctx;
  const bool rmv = rmv_due(s), adj = adj_due(s);
  if (!rmv && !adj) return 0;
  RET(mlg_synchronize(c));
  if (rmv) {
    RET(mlg_remove(psi, chi, b));
    RET(c->g.p.adams ? rich_fe_be(s, psi, chi, b) : rich_etdfe_be(s, psi, chi, b));
    RET(launch_diagnost(c, psi->e));
    RET(mlg_synchronize(c));
  }
  if (adj) RET(mlg_hyp3adj(c, psi, chi, b));
  return 0;
}
// steps before the next event, at most `left`
static int steps_to_event(const mlg_solver *s, int left) {
  int k = left;
  if (s->rmv_sw != 0 && s->rmv_int > 0) k = std::min(k, s->rmv_int - s->n % s->rmv_int);
  if (s->adj_sw != 0 && s->adj_int > 0) k = std::min(k, s->adj_int - s->n % s->adj_int);
  return k;
}

extern "C" int mlg_time_step(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b) {
  if (!s || !psi || !chi || !b) return MLG_E_ARG;
  RET(one_step(s, psi, chi, b));
  RET(march_events(s, psi, chi, b));
  return mlg_synchronize(s->ctx);
}

extern "C" int mlg_time_steps(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps) {
  if (!s || !psi || !chi || !b || nsteps < 0) return MLG_E_ARG;
  for (int done = 0; done < nsteps;) {
    const int k = steps_to_event(s, nsteps - done);
    for (int i = 0; i < k; i++) RET(one_step(s, psi, chi, b));
    done += k;
    RET(march_events(s, psi, chi, b));
  }
  return mlg_synchronize(s->ctx);
}

// The graph's temporaries go back to the pool and its factor sets are
// unpinned (after the queued replays have finished with them).
static void drop_graph(mlg_solver *s) {
  mlg_ctx *c = s->ctx;
  if (s->gexec || !s->gpool.empty()) (void)hipStreamSynchronize(c->stream);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  s->gpsi = s->gchi = s->gb = nullptr;
  for (auto &a : s->gaddr) a = nullptr;
  for (mlg_field *f : s->gpool) c->pool.push_back(f);
  s->gpool.clear();
  for (double *p : s->gfactors)
    for (auto &f : c->fcache)
      if (f.f == p && f.pins > 0) f.pins--;
  s->gfactors.clear();
}

static void field_addrs(const mlg_field *psi, const mlg_field *chi, const mlg_field *b, const void **a) {
  a[0] = psi->e, a[1] = chi->e, a[2] = b->e, a[3] = psi->ln, a[4] = chi->ln, a[5] = b->ln;
}

// One step captured into a hipGraph and replayed.  A step keeps psi, chi, b
// and the solver's history fields in their buffers (transforms only swap
// buffers of temporaries), so one captured step is valid for every later
// step on the same fields while their device buffers stay the same (a
// transform of psi, chi or b between calls may swap them: then the step is
// captured again).  The eager step that precedes capture fills the temporary
// pool and the band-factor cache, so the captured step allocates and clears
// nothing.  After capture the graph owns the pooled temporaries (they leave
// the pool, so no later call can take one over as its own storage) and pins
// the factor sets in the cache.
static int steps_graph(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps);
extern "C" int mlg_time_steps_graph(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps) {
  if (!s || !psi || !chi || !b || nsteps < 0) return MLG_E_ARG;
  for (int done = 0; done < nsteps;) {
    const int k = steps_to_event(s, nsteps - done);
    RET(steps_graph(s, psi, chi, b, k));
    done += k;
    RET(march_events(s, psi, chi, b));
  }
  return 0;
}

// the captured steps between two events: an eager step and the capture
// when the graph is missing or stale, then replays
static int steps_graph(mlg_solver *s, mlg_field *psi, mlg_field *chi, mlg_field *b, int nsteps) {
  if (nsteps == 0) return 0;
  mlg_ctx *c = s->ctx;
  int done = 0;
  const void *addr[6];
  field_addrs(psi, chi, b, addr);
  bool same_buffers = true;
  for (int i = 0; i < 6; i++) same_buffers = same_buffers && addr[i] == s->gaddr[i];
  const int split_key = c->leg_split ? c->split_fp64_modes : 0;
  if (!s->gexec || s->gpsi != psi || s->gchi != chi || s->gb != b || !same_buffers || s->gnrchop != c->g.nrchop ||
      s->gstream != c->stream || s->gsplit != split_key) {
    drop_graph(s);
    RET(one_step(s, psi, chi, b));
    done = 1;
    RET(mlg_synchronize(c));
    if (done == nsteps) return 0;
    const bool timer = c->timer.on;
    c->timer.on = false;
    const double t0 = s->t;
    const int n0 = s->n;
    MLG_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = one_step(s, psi, chi, b);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &gr);
    c->timer.on = timer;
    s->t = t0;
    s->n = n0;
    if (rc || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc ? rc : MLG_E_HIP;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(gr);
      return MLG_E_HIP;
    }
    s->graph = gr;
    s->gexec = ex;
    s->gpsi = psi;
    s->gchi = chi;
    s->gb = b;
    field_addrs(psi, chi, b, s->gaddr);
    s->gpool.swap(c->pool);  // every temporary the captured step used is in the pool now
    for (auto &f : c->fcache) {
      f.pins++;
      s->gfactors.push_back(f.f);
    }
    s->gnrchop = c->g.nrchop;
    s->gsplit = split_key;
    s->gstream = c->stream;
  }
  for (int i = done; i < nsteps; i++) {
    MLG_CHECK_HIP(hipGraphLaunch(s->gexec, c->stream));
    s->t += c->g.p.dt;
    s->n += 1;
  }
  return mlg_synchronize(c);
}

// ------------------------------------------------------------------ initial condition
// QVORTEX (src/mod_layout.f90:449-610) with LAY (:64-118) and GAUS (:640-673)
extern "C" int mlg_qvortex(mlg_ctx *c, double q, mlg_field *psi, mlg_field *chi) {
  if (!c || !psi || !chi) return MLG_E_ARG;
  const Grid &G = c->g;
  const double pi = std::acos(-1.0);
  std::vector<double2> h((size_t)c->n_ppp, make_double2(0, 0));
  std::vector<double> thr(G.nth), thi(G.nth);
  for (int j = 0; j < G.nth; j++) {
    thr[j] = 2 * pi / G.p.minc / G.nth * j;
    thi[j] = 2 * pi / G.p.minc / G.nth * (j + 0.5);
  }
  for (int z = 0; z < G.ndimx; z++)
    for (int j = 0; j < G.nth; j++)
      for (int r = 0; r < G.nr; r++) {
        double rr = G.r[r];
        double f = std::pow(1 - G.x[r], 2);
        auto gaus = [&](double th) {
          double xx = rr * std::cos(th), yy = rr * std::sin(th);
          double r2 = (std::pow(xx, 2.0) + std::pow(yy, 2.0)) / std::pow(1.0, 2.0);
          return (q * 2.0) / std::pow(1.0, 2.0) * std::exp(-1.0 * r2);
        };
        h[(size_t)r + (size_t)G.ndimr * ((size_t)j + (size_t)G.ndimth * z)] = make_double2(gaus(thr[j]) / f, gaus(thi[j]) / f);
      }
  Temps T(c);
  mlg_field *a = T.get(), *bb = T.get();
  RET(T.ready());
  RET(mlg_field_upload(a, (const double *)h.data(), MLG_PPP, 0.0));
  RET(mlg_toff(a));
  RET(launch_negate(c, a->e, c->n_fff));
  RET(launch_set_ln(c, a->ln, nullptr, 0.0, 0.0));
  RET(mlg_idelsqh(a, bb));
  RET(copy0(psi, bb));
  RET(mlg_field_zero(chi, MLG_FFF));
  return mlg_synchronize(c);
}
