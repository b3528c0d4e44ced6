// Batched sweep of independent runs of different resolution, and the tile
// tuner of the grouped RTRAN backward products.
//
// The reference runs one case per job (SOLVER_T%TIME_STEPPING,
// src/mod_march.f90:244-300); a parameter sweep there is one job per case.
// Here all runs of a sweep step together in one hipGraph launch per step.
// Each run's step is captured on its own context's stream and cut into
// segments at its RTRAN backward products (launch_rtran_bwd_n defers them
// while the context records); the products found at the same place of every
// run's step then run as ONE grouped persistent launch (k_rtran_bwd_grouped)
// over (run, mode bucket) jobs.  Runs share no data, so segments of
// different runs have no edges between them; a grouped launch waits for the
// segments before it of every run it serves, and those runs' next segments
// wait for it.
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "mlegs_internal.h"

using namespace mlg;

#define RET(x)                 \
  do {                         \
    int _r = (x);              \
    if (_r) return _r;         \
  } while (0)

namespace mlg {

struct SweepRec {
  std::vector<hipGraph_t> segs;  // segment p runs before deferred product p
  std::vector<BwdJob> bwd;       // the deferred products, all modes, default shape
};

int sweep_defer_bwd(mlg_ctx *c, const KGrid &g, const TransformJobs &jobs, int nj) {
  SweepRec *r = c->rec;
  hipGraph_t seg = nullptr;
  MLG_CHECK_HIP(hipStreamEndCapture(c->stream, &seg));
  r->segs.push_back(seg);
  BwdJob j;
  std::memset(&j, 0, sizeof(j));
  j.g = g;
  j.pf = c->d.pf;
  j.lnx = c->d.lnx;
  j.jobs = jobs;
  j.nj = nj;
  j.shape = BWD_TILE_64x64;
  r->bwd.push_back(j);
  MLG_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

// ---------------------------------------------------------------- tuner
// Mode buckets by the K extent NRCHOPS(m) of the product (the number of
// 32-row K chunks a tile runs): 0: K > 128 (5 chunks or more), 1: 65..128
// (3-4), 2: 33..64 (2), 3: K <= 32 (1; and modes past the truncation).
constexpr int BWD_BUCKETS = 4;
static int bwd_bucket(int k) { return k > 128 ? 0 : k > 64 ? 1 : k > 32 ? 2 : 3; }

// The job's modes split by bucket (K falls as m rises, so each bucket is a
// contiguous range of modes); empty buckets are left out.
static void bucket_jobs(const BwdJob &j, std::vector<std::pair<int, BwdJob>> &out) {
  const KGrid &g = j.g;
  int m = g.mlo;
  const int mend = g.mlo + g.mcnt;
  while (m < mend) {
    const int b = bwd_bucket(nrchops_of(g.nrchop, m * g.minc));
    int e = m + 1;
    while (e < mend && bwd_bucket(nrchops_of(g.nrchop, e * g.minc)) == b) e++;
    BwdJob s = j;
    s.g.mlo = m;
    s.g.mcnt = e - m;
    out.emplace_back(b, s);
    m = e;
  }
}

static std::string tune_key(const BwdJob &j) {
  const KGrid &g = j.g;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d", g.nr, g.nrh, g.ndimr, g.npf, g.nrchopdim,
                g.ntchopdim, g.nxchopdim, g.nrchop, g.minc, g.mlo, g.mcnt, j.nj);
  return buf;
}

struct TuneEntry {
  int shape;
  std::array<double, BWD_TILE_COUNT> us;
};
static std::mutex tune_mu;
static std::map<std::string, TuneEntry> tune_table;

static int device_cus(int dev) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
  return n;
}

// Times the job (its modes, nj fields) with every tile shape in the grouped
// kernel on the context's stream, on scratch inputs (zeros) and outputs.
static int time_bwd_job(mlg_ctx *c, const BwdJob &job, TuneEntry &te) {
  const KGrid &g = job.g;
  const size_t nin = (size_t)g.nrchopdim * g.ntchopdim * g.nxchopdim;
  const size_t nout = (size_t)g.ndimr * g.ntchopdim * g.nxchopdim;
  const size_t bytes = ((nin + nout) * job.nj) * sizeof(double2) + 8 * sizeof(double);
  char *buf = nullptr;
  BwdJob *dj = nullptr;
  int *db0 = nullptr;
  hipEvent_t ev[12] = {};
  int rc = 0;
  if (hipMalloc(&buf, bytes) != hipSuccess) return MLG_E_NOMEM;
  if (hipMalloc(&dj, sizeof(BwdJob)) != hipSuccess || hipMalloc(&db0, 2 * sizeof(int)) != hipSuccess) rc = MLG_E_NOMEM;
  for (int i = 0; i < 12 && !rc; i++)
    if (hipEventCreate(&ev[i]) != hipSuccess) rc = MLG_E_HIP;
  if (!rc && hipMemsetAsync(buf, 0, bytes, c->stream) != hipSuccess) rc = MLG_E_HIP;
  const int cus = device_cus(c->device);
  for (int sh = 0; sh < BWD_TILE_COUNT && !rc; sh++) {
    BwdJob j = job;
    double2 *in0 = (double2 *)buf, *out0 = in0 + nin * job.nj;
    const double *ln0 = (const double *)(out0 + nout * job.nj);
    for (int f = 0; f < job.nj; f++) {
      j.jobs.in[f] = in0 + nin * f;
      j.jobs.ln[f] = ln0;
      j.jobs.out[f] = out0 + nout * f;
    }
    j.shape = sh;
    const int total = bwd_job_tiles(j);
    const int b0[2] = {0, total};
    if (hipMemcpyAsync(dj, &j, sizeof(j), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(db0, b0, sizeof(b0), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      rc = MLG_E_HIP;
      break;
    }
    const int grid = bwd_grouped_grid(total, cus);
    for (int w = 0; w < 2 && !rc; w++) rc = launch_rtran_bwd_grouped(c->stream, dj, db0, 1, total, grid);
    for (int i = 0; i < 5 && !rc; i++) {
      (void)hipEventRecord(ev[2 * i], c->stream);
      rc = launch_rtran_bwd_grouped(c->stream, dj, db0, 1, total, grid);
      (void)hipEventRecord(ev[2 * i + 1], c->stream);
    }
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = MLG_E_HIP;
    float t[5];
    for (int i = 0; i < 5 && !rc; i++)
      if (hipEventElapsedTime(&t[i], ev[2 * i], ev[2 * i + 1]) != hipSuccess) rc = MLG_E_HIP;
    if (!rc) {
      std::sort(t, t + 5);
      te.us[sh] = 1e3 * t[2];
    }
  }
  if (!rc) {
    te.shape = 0;
    for (int sh = 1; sh < BWD_TILE_COUNT; sh++)
      if (te.us[sh] < te.us[te.shape]) te.shape = sh;
  }
  (void)hipStreamSynchronize(c->stream);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (db0) (void)hipFree(db0);
  if (dj) (void)hipFree(dj);
  (void)hipFree(buf);
  return rc;
}

// The tuned shape of a bucket job (timed on first use, then kept).
static int tuned_shape(mlg_ctx *c, const BwdJob &j, int *shape, bool retime = false, TuneEntry *out = nullptr) {
  const std::string key = tune_key(j);
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_table.find(key);
    if (it != tune_table.end() && !retime) {
      *shape = it->second.shape;
      if (out) *out = it->second;
      return 0;
    }
  }
  TuneEntry te;
  te.us.fill(0.0);
  RET(time_bwd_job(c, j, te));
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_table[key] = te;
  }
  *shape = te.shape;
  if (out) *out = te;
  return 0;
}

}  // namespace mlg

extern "C" int mlg_tune_rtran_bwd(mlg_ctx *c, int nfields, int shape[4], double us[16]) {
  if (!c || nfields < 1 || nfields > MAX_TRANSFORM_JOBS || !shape || !us) return MLG_E_ARG;
  RET(mlg_synchronize(c));
  BwdJob j;
  std::memset(&j, 0, sizeof(j));
  j.g = kgrid(c);
  j.pf = c->d.pf;
  j.lnx = c->d.lnx;
  j.nj = nfields;
  std::vector<std::pair<int, BwdJob>> parts;
  bucket_jobs(j, parts);
  for (int b = 0; b < BWD_BUCKETS; b++) shape[b] = -1;
  for (int i = 0; i < 16; i++) us[i] = 0.0;
  for (auto &p : parts) {
    TuneEntry te;
    RET(tuned_shape(c, p.second, &shape[p.first], true, &te));
    for (int s = 0; s < BWD_TILE_COUNT; s++) us[4 * p.first + s] = te.us[s];
  }
  return 0;
}

// ---------------------------------------------------------------- sweep
struct mlg_sweep {
  int nruns = 0, flags = 0;
  std::vector<mlg_solver *> s;
  std::vector<mlg_field *> psi, chi, b;
  std::vector<std::array<const void *, 6>> addr;  // field buffers at record
  std::vector<int> nrchop;                         // truncation at record
  std::vector<std::vector<mlg_field *>> pools;     // temporaries the graph uses
  std::vector<std::vector<double *>> factors;      // band factors the graph reads
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;  // the sweep's graph (one launch per step)
  void *dmem = nullptr;           // job tables of the grouped launches
  // MLG_SWEEP_STREAMS: each run's segments as graphs of their own, launched
  // on the run's stream; the grouped launches on the sweep's stream, joined
  // to the runs they serve by events
  struct Grouped {
    const BwdJob *dj;
    const int *db0;
    int njobs, total, grid;
    std::vector<int> served;
    int pos;  // the place in the step (the runs' p-th backward product)
  };
  std::vector<Grouped> gl;
  std::vector<std::vector<hipGraphExec_t>> segx;  // [run][segment] (null: empty)
  std::vector<std::vector<hipEvent_t>> segev;     // [run][position]: after the segment before it
  std::vector<hipEvent_t> gev;                    // [launch]: after the grouped launch
  std::vector<std::vector<int>> pos_launch;       // [position]: its grouped launches
  bool built = false;
  double gflop = 0.0;  // algorithmic GFLOP of the grouped launches of one step
  int info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<int> err;  // each run's first DIAGNOST code (0: none)
};

// Reads every run's sticky device code: DIAGNOST codes (below MLG_E_ARG)
// are kept per run, the first library/HIP error is returned.
static int collect_errors(mlg_sweep *w) {
  int lib = 0;
  for (int r = 0; r < w->nruns; r++) {
    const int e = mlg_synchronize(w->s[r]->ctx);
    if (e >= MLG_E_ARG) {
      if (!lib) lib = e;
    } else if (e && !w->err[r]) {
      w->err[r] = e;
    }
  }
  return lib;
}

// The DIAGNOST code of the lowest-numbered run that has one (0: none).
static int first_code(const mlg_sweep *w) {
  for (int e : w->err)
    if (e) return e;
  return 0;
}

static std::array<const void *, 6> run_addrs(const mlg_sweep *w, int r) {
  return {w->psi[r]->e, w->chi[r]->e, w->b[r]->e, w->psi[r]->ln, w->chi[r]->ln, w->b[r]->ln};
}

// The graph goes; its temporaries return to their pools and its factor
// sets are unpinned.
static void sweep_drop(mlg_sweep *w) {
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  for (int r = 0; r < (int)w->segx.size(); r++)
    if (w->s[r]->ctx->stream) (void)hipStreamSynchronize(w->s[r]->ctx->stream);
  if (w->exec) (void)hipGraphExecDestroy(w->exec);
  w->exec = nullptr;
  for (auto &v : w->segx)
    for (hipGraphExec_t x : v)
      if (x) (void)hipGraphExecDestroy(x);
  w->segx.clear();
  for (auto &v : w->segev)
    for (hipEvent_t e : v)
      if (e) (void)hipEventDestroy(e);
  w->segev.clear();
  for (hipEvent_t e : w->gev)
    if (e) (void)hipEventDestroy(e);
  w->gev.clear();
  w->gl.clear();
  w->pos_launch.clear();
  w->built = false;
  w->gflop = 0.0;
  if (w->dmem) (void)hipFree(w->dmem);
  w->dmem = nullptr;
  for (int r = 0; r < (int)w->pools.size(); r++) {
    mlg_ctx *c = w->s[r]->ctx;
    for (mlg_field *f : w->pools[r]) c->pool.push_back(f);
    w->pools[r].clear();
    for (double *p : w->factors[r])
      for (auto &f : c->fcache)
        if (f.f == p && f.pins > 0) f.pins--;
    w->factors[r].clear();
  }
}

static void free_recs(std::vector<SweepRec> &rec) {
  for (auto &r : rec)
    for (hipGraph_t g : r.segs)
      if (g) (void)hipGraphDestroy(g);
  rec.clear();
}

// Records every run's step and builds the sweep's graph.
static int sweep_build(mlg_sweep *w) {
  const bool grouped = (w->flags & MLG_SWEEP_GROUPED) != 0;
  const int forced = (w->flags >> 4) & 7;  // 0: default or tuned shape, else shape + 1
  std::vector<SweepRec> rec(w->nruns);
  int rc = 0;
  for (int r = 0; r < w->nruns && !rc; r++) {
    mlg_solver *s = w->s[r];
    mlg_ctx *c = s->ctx;
    const bool timer = c->timer.on;
    const double t0 = s->t;
    const int n0 = s->n;
    c->timer.on = false;
    if (grouped) c->rec = &rec[r];
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      rc = step_once(s, w->psi[r], w->chi[r], w->b[r]);
      hipGraph_t last = nullptr;
      e = hipStreamEndCapture(c->stream, &last);
      if (last) rec[r].segs.push_back(last);
    }
    c->rec = nullptr;
    c->timer.on = timer;
    s->t = t0;
    s->n = n0;
    if (!rc && e != hipSuccess) rc = MLG_E_HIP;
    if (rc) break;
    // the graph owns every pooled temporary and pins every factor set
    w->pools[r].swap(c->pool);
    for (auto &f : c->fcache) {
      f.pins++;
      w->factors[r].push_back(f.f);
    }
    w->addr[r] = run_addrs(w, r);
    w->nrchop[r] = c->g.nrchop;
  }
  if (rc) {
    free_recs(rec);
    sweep_drop(w);
    return rc;
  }
  // grouped positions: position p serves the runs with a p-th product; with
  // MLG_SWEEP_BY_GRID each class of runs of one grid has launches of its own
  int npos = 0;
  for (auto &r : rec) npos = std::max(npos, (int)r.bwd.size());
  std::vector<int> cls(w->nruns, 0);
  int ncls = 1;
  if (w->flags & MLG_SWEEP_BY_GRID) {
    std::map<std::array<int, 6>, int> ids;
    for (int r = 0; r < w->nruns; r++) {
      const Grid &G = w->s[r]->ctx->g;
      const std::array<int, 6> key = {G.nr, G.nth, G.nx, G.nrchop, G.ntchop, G.nxchop};
      auto it = ids.find(key);
      if (it == ids.end()) {
        const int id = (int)ids.size();
        ids[key] = id;
        cls[r] = id;
      } else {
        cls[r] = it->second;
      }
    }
    ncls = (int)ids.size();
  }
  const int nlaunch = npos * ncls;  // launch l = position l / ncls, class l % ncls
  std::vector<std::vector<BwdJob>> pj(nlaunch);
  std::vector<std::vector<int>> pb(nlaunch);
  std::vector<std::vector<int>> served(nlaunch);
  int shape_count[BWD_TILE_COUNT] = {};
  for (int l = 0; l < nlaunch && !rc; l++) {
    const int p = l / ncls;
    int total = 0;
    for (int r = 0; r < w->nruns && !rc; r++) {
      if (p >= (int)rec[r].bwd.size() || cls[r] != l % ncls) continue;
      served[l].push_back(r);
      std::vector<std::pair<int, BwdJob>> parts;
      bucket_jobs(rec[r].bwd[p], parts);
      for (auto &q : parts) {
        BwdJob j = q.second;
        if (forced) j.shape = std::min(forced - 1, BWD_TILE_COUNT - 1);
        else if (w->flags & MLG_SWEEP_TUNE) rc = tuned_shape(w->s[r]->ctx, j, &j.shape);
        if (rc) break;
        shape_count[j.shape]++;
        pb[l].push_back(total);
        total += bwd_job_tiles(j);
        pj[l].push_back(j);
      }
    }
    pb[l].push_back(total);
  }
  // one device allocation for every launch's tables
  std::vector<size_t> off_j(nlaunch), off_b(nlaunch);
  size_t bytes = 0;
  for (int l = 0; l < nlaunch; l++) {
    off_j[l] = bytes;
    bytes += pj[l].size() * sizeof(BwdJob);
    bytes = (bytes + 255) & ~(size_t)255;
    off_b[l] = bytes;
    bytes += pb[l].size() * sizeof(int);
    bytes = (bytes + 255) & ~(size_t)255;
  }
  if (!rc && bytes) {
    if (hipMalloc(&w->dmem, bytes) != hipSuccess) rc = MLG_E_NOMEM;
    for (int l = 0; l < nlaunch && !rc; l++)
      if (hipMemcpy((char *)w->dmem + off_j[l], pj[l].data(), pj[l].size() * sizeof(BwdJob), hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy((char *)w->dmem + off_b[l], pb[l].data(), pb[l].size() * sizeof(int), hipMemcpyHostToDevice) !=
              hipSuccess)
        rc = MLG_E_HIP;
  }
  // the grouped launches (a class with no run at a position has none)
  const int cus = device_cus(w->s[0]->ctx->device);
  int gjobs = 0;
  std::vector<std::vector<int>> at_pos(npos);  // launches of each position
  for (int l = 0; l < nlaunch && !rc; l++) {
    if (served[l].empty()) continue;
    mlg_sweep::Grouped q;
    q.dj = (const BwdJob *)((char *)w->dmem + off_j[l]);
    q.db0 = (const int *)((char *)w->dmem + off_b[l]);
    q.njobs = (int)pj[l].size();
    q.total = pb[l].back();
    q.grid = bwd_grouped_grid(q.total, cus);
    q.served = served[l];
    q.pos = l / ncls;
    for (const BwdJob &j : pj[l]) {
      long sk = 0;
      for (int m = j.g.mlo; m < j.g.mlo + j.g.mcnt; m++) sk += nrchops_of(j.g.nrchop, m * j.g.minc);
      w->gflop += 2.0 * j.g.nr * j.g.nxchopdim * j.nj * (double)sk * 1e-9;
    }
    gjobs += q.njobs;
    at_pos[q.pos].push_back((int)w->gl.size());
    w->gl.push_back(q);
  }
  int nodes = 0;
  if (!rc && (w->flags & MLG_SWEEP_STREAMS)) {
    w->segx.assign(w->nruns, {});
    w->segev.assign(w->nruns, std::vector<hipEvent_t>(npos, nullptr));
    w->gev.assign(w->gl.size(), nullptr);
    for (int r = 0; r < w->nruns && !rc; r++)
      for (hipGraph_t g : rec[r].segs) {
        size_t n = 0;
        hipGraphExec_t x = nullptr;
        if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess ||
            (n > 0 && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess))
          rc = MLG_E_HIP;
        w->segx[r].push_back(x);
        nodes += n > 0;
      }
    for (int l = 0; l < (int)w->gl.size() && !rc; l++) {
      if (hipEventCreateWithFlags(&w->gev[l], hipEventDisableTiming) != hipSuccess) rc = MLG_E_HIP;
      for (int r : w->gl[l].served)
        if (!rc && hipEventCreateWithFlags(&w->segev[r][w->gl[l].pos], hipEventDisableTiming) != hipSuccess)
          rc = MLG_E_HIP;
    }
  } else if (!rc) {
    // one graph: each run's segments in order, a grouped launch after the
    // segments before it of every run it serves
    hipGraph_t G = nullptr;
    if (hipGraphCreate(&G, 0) != hipSuccess) rc = MLG_E_HIP;
    std::vector<hipGraphNode_t> last(w->nruns, nullptr);
    for (int p = 0; p <= npos && !rc; p++) {
      for (int r = 0; r < w->nruns && !rc; r++) {
        if (p >= (int)rec[r].segs.size()) continue;
        size_t n = 0;
        if (hipGraphGetNodes(rec[r].segs[p], nullptr, &n) != hipSuccess) rc = MLG_E_HIP;
        if (rc || n == 0) continue;
        hipGraphNode_t node = nullptr;
        if (hipGraphAddChildGraphNode(&node, G, last[r] ? &last[r] : nullptr, last[r] ? 1 : 0, rec[r].segs[p]) !=
            hipSuccess)
          rc = MLG_E_HIP;
        last[r] = node;
        nodes++;
      }
      if (p == npos || rc) continue;
      for (int l : at_pos[p]) {
        mlg_sweep::Grouped &q = w->gl[l];
        std::vector<hipGraphNode_t> deps;
        for (int r : q.served)
          if (last[r]) deps.push_back(last[r]);
        void *args[4] = {(void *)&q.dj, (void *)&q.db0, (void *)&q.njobs, (void *)&q.total};
        hipKernelNodeParams kp;
        std::memset(&kp, 0, sizeof(kp));
        kp.func = bwd_grouped_kernel();
        kp.gridDim = dim3(q.grid);
        kp.blockDim = dim3(256);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        if (hipGraphAddKernelNode(&node, G, deps.data(), deps.size(), &kp) != hipSuccess) rc = MLG_E_HIP;
        for (int r : q.served) last[r] = node;
        nodes++;
        if (rc) break;
      }
    }
    if (!rc && hipGraphInstantiate(&w->exec, G, nullptr, nullptr, 0) != hipSuccess) rc = MLG_E_HIP;
    if (G) (void)hipGraphDestroy(G);
  }
  free_recs(rec);
  if (rc) {
    sweep_drop(w);
    return rc;
  }
  w->built = true;
  w->pos_launch = at_pos;
  w->info[0] = npos;
  w->info[1] = nodes;
  w->info[2] = gjobs;
  for (int s = 0; s < BWD_TILE_COUNT; s++) w->info[3 + s] = shape_count[s];
  w->info[6] = (int)w->gl.size();
  w->info[7] = w->nruns;
  return 0;
}

extern "C" int mlg_sweep_create(int nruns, mlg_solver *const *solvers, mlg_field *const *psi, mlg_field *const *chi,
                                mlg_field *const *b, int flags, mlg_sweep **out) {
  if (!out || nruns < 1 || !solvers || !psi || !chi || !b || (flags & ~0xff)) return MLG_E_ARG;
  // per-grid launches only in streams mode: built as one graph (grouped
  // kernel nodes in independent branches), replays gave results different
  // from the runs stepped alone, with no cause found in the graph's
  // construction; that form is not offered
  if ((flags & MLG_SWEEP_BY_GRID) && !(flags & MLG_SWEEP_STREAMS)) return MLG_E_ARG;
  *out = nullptr;
  const int dev = solvers[0] ? (solvers[0]->ctx ? solvers[0]->ctx->device : -1) : -1;
  for (int r = 0; r < nruns; r++) {
    mlg_solver *s = solvers[r];
    if (!s || !s->ctx || !psi[r] || !chi[r] || !b[r]) return MLG_E_ARG;
    mlg_ctx *c = s->ctx;
    if (c->device != dev || psi[r]->ctx != c || chi[r]->ctx != c || b[r]->ctx != c) return MLG_E_ARG;
    if ((flags & MLG_SWEEP_GROUPED) && c->leg_split) return MLG_E_ARG;
    // the sweep steps without TIME_STEPPING's REMOVE / HYP3ADJ events
    if ((s->rmv_sw != 0 && s->rmv_int > 0) || (s->adj_sw != 0 && s->adj_int > 0)) return MLG_E_ARG;
    for (int q = 0; q < r; q++)
      if (solvers[q] == s || solvers[q]->ctx == c) return MLG_E_ARG;  // one run per context
  }
  mlg_sweep *w = new mlg_sweep();
  w->nruns = nruns;
  w->flags = flags;
  w->s.assign(solvers, solvers + nruns);
  w->psi.assign(psi, psi + nruns);
  w->chi.assign(chi, chi + nruns);
  w->b.assign(b, b + nruns);
  w->addr.resize(nruns);
  w->nrchop.assign(nruns, -1);
  w->pools.resize(nruns);
  w->factors.resize(nruns);
  w->err.assign(nruns, 0);
  if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
    delete w;
    return MLG_E_HIP;
  }
  *out = w;
  return 0;
}

extern "C" int mlg_sweep_steps(mlg_sweep *w, int nsteps) {
  if (!w || nsteps < 0) return MLG_E_ARG;
  if (nsteps == 0) return 0;
  for (int r = 0; r < w->nruns; r++)
    if (w->psi[r]->space != MLG_FFF || w->chi[r]->space != MLG_FFF || w->b[r]->space != MLG_FFF) return MLG_E_SPACE;
  bool stale = !w->built;
  for (int r = 0; r < w->nruns && !stale; r++)
    stale = run_addrs(w, r) != w->addr[r] || w->nrchop[r] != w->s[r]->ctx->g.nrchop;
  int done = 0;
  if (stale) {
    sweep_drop(w);
    // one eager step of every run (each on its own stream) fills the
    // temporary pools and the band-factor caches the recorded step uses
    for (int r = 0; r < w->nruns; r++) RET(step_once(w->s[r], w->psi[r], w->chi[r], w->b[r]));
    RET(collect_errors(w));
    done = 1;
    if (done == nsteps) return first_code(w);
    RET(sweep_build(w));
  }
  for (int r = 0; r < w->nruns; r++) MLG_CHECK_HIP(hipStreamSynchronize(w->s[r]->ctx->stream));
  const bool streams = (w->flags & MLG_SWEEP_STREAMS) != 0;
  for (int i = done; i < nsteps; i++) {
    if (!streams) {
      MLG_CHECK_HIP(hipGraphLaunch(w->exec, w->stream));
    } else {
      const int npos = (int)w->pos_launch.size();
      for (int p = 0; p <= npos; p++) {
        for (int r = 0; r < w->nruns; r++)
          if (p < (int)w->segx[r].size() && w->segx[r][p])
            MLG_CHECK_HIP(hipGraphLaunch(w->segx[r][p], w->s[r]->ctx->stream));
        if (p == npos) break;
        for (int l : w->pos_launch[p]) {
          const mlg_sweep::Grouped &q = w->gl[l];
          for (int r : q.served) {
            MLG_CHECK_HIP(hipEventRecord(w->segev[r][p], w->s[r]->ctx->stream));
            MLG_CHECK_HIP(hipStreamWaitEvent(w->stream, w->segev[r][p], 0));
          }
          RET(launch_rtran_bwd_grouped(w->stream, q.dj, q.db0, q.njobs, q.total, q.grid));
          MLG_CHECK_HIP(hipEventRecord(w->gev[l], w->stream));
          for (int r : q.served) MLG_CHECK_HIP(hipStreamWaitEvent(w->s[r]->ctx->stream, w->gev[l], 0));
        }
      }
    }
    for (mlg_solver *s : w->s) {
      s->t += s->ctx->g.p.dt;
      s->n += 1;
    }
  }
  MLG_CHECK_HIP(hipStreamSynchronize(w->stream));
  for (int r = 0; r < w->nruns && streams; r++) MLG_CHECK_HIP(hipStreamSynchronize(w->s[r]->ctx->stream));
  RET(collect_errors(w));
  return first_code(w);
}

extern "C" int mlg_sweep_work(const mlg_sweep *w, double *gflop_per_step) {
  if (!w || !gflop_per_step) return MLG_E_ARG;
  *gflop_per_step = w->gflop;
  return 0;
}

extern "C" int mlg_sweep_codes(const mlg_sweep *w, int *codes) {
  if (!w || !codes) return MLG_E_ARG;
  for (int r = 0; r < w->nruns; r++) codes[r] = w->err[r];
  return 0;
}

extern "C" int mlg_sweep_info(const mlg_sweep *w, int out[8]) {
  if (!w || !out) return MLG_E_ARG;
  for (int i = 0; i < 8; i++) out[i] = w->info[i];
  return 0;
}

extern "C" int mlg_sweep_destroy(mlg_sweep *w) {
  if (!w) return 0;
  sweep_drop(w);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
  return 0;
}
