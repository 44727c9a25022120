// Batched sweep of many independent Taylor-Couette instances on one MI355X.
//
// Each instance is an ordinary tc_ctx (its own shape, parameters, fields and
// LU tables).  Instances with the same grid shape and schedule form a group;
// every kernel of the fused RK3 step then runs as one grouped launch per
// group, blockIdx selecting the instance through a device table of
// StepDesc.  Groups run on their own streams, forked from and joined back to
// the batch stream, and one whole step of all groups can be captured as a
// hipGraph and replayed.  Per-step scalars (inner-wall oscillation terms,
// the logging flag, the step index) reach the device through a small
// pinned-memory copy before each step; device-detected exit conditions
// accumulate in per-instance BatchRes records that the host reads every
// sync_every steps.
//
// The per-instance arithmetic is exactly that of tc_step with the fused
// kernels, so an instance's fields do not depend on which other instances
// share its batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "tc_internal.h"

using tc::BatchDyn;
using tc::BatchRes;
using tc::StepDesc;

namespace {

struct Inst {
  tc_ctx *c = nullptr;
  bool live = false;   // slot in use
  bool done = false;   // finished: code holds the exit code
  int steps = 0;       // steps taken = ntime of the next step
  double time = 0.0;   // host copy of time (advanced exactly as tc_step does)
  int code = 0;
  int end_step = -1;   // ntime of the step that ended the run
  int host_end = -1;   // ntime of a host-detected end (333/555), pending harvest
  int host_code = 0;
  double out9[9] = {0};
  int group = -1, slot = -1;
};

struct Group {
  std::vector<int> ids;
  tc_ctx *rep = nullptr;
  StepDesc *tab[3] = {nullptr, nullptr, nullptr};
  BatchDyn *dyn = nullptr;
  BatchRes *res = nullptr;
  std::vector<BatchRes> res_host;
  hipStream_t stream = nullptr;
  hipEvent_t join = nullptr;
  size_t dyn_off = 0;  // offset of this group's entries in the staging slot
};

using Key = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int, int>;

Key group_key(const tc_ctx *c) {
  return Key(c->nthm, c->nzm, c->nrm, c->p.nsst, c->helm_lpt, c->poisson_pq,
             c->masks_uniform_one ? 1 : 0, c->plan_t.lines, c->plan_t.me, c->plan_t.padsh,
             c->plan_z.lines, c->plan_z.me * 1000 + c->plan_z.nt, c->plan_z.padsh,
             tc::defers_correction(c) ? 1 : 0);
}

// a failed call is also cleared from the runtime's last-error state
#define BCHK(x)                                                                  \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      (void)hipGetLastError();                                                   \
      return tc::set_error(std::string(#x) + ": " + hipGetErrorString(e_));      \
    }                                                                            \
  } while (0)

}  // namespace

struct tc_batch {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<Inst> inst;
  std::vector<Group> groups;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  static constexpr int RING = 4;
  BatchDyn *dyn_host[RING] = {nullptr};
  hipEvent_t dyn_ev[RING] = {nullptr};
  size_t dyn_cap = 0;
  int ring = 0;
  int launches = 0;
};

namespace {

void free_groups(tc_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (Group &g : b->groups) {
    for (int s = 0; s < 3; s++)
      if (g.tab[s]) hipFree(g.tab[s]);
    if (g.dyn) hipFree(g.dyn);
    if (g.res) hipFree(g.res);
    if (g.join) hipEventDestroy(g.join);
    if (g.stream) hipStreamDestroy(g.stream);
  }
  b->groups.clear();
}

// out9 of the last step an instance took (its end step once it is done),
// as tc_step reports it: the fixed-dt CFL (cfl * dt, main.F90:172,201) of
// that step's logging pass or, for a step without one, of its start; the
// diagnostics of the last logging pass; dt and the time after the step.
void set_out9(Inst &I, const BatchRes &r) {
  const tc_ctx *c = I.c;
  const int last = I.done ? I.end_step : I.steps - 1;
  if (last < 0) return;
  const double dt = c->p.dtmax;
  if (r.log_ntime >= 0) {
    I.out9[3] = r.diag[1];
    tc::ctx_torque(c, r.diag[8], r.diag[9], &I.out9[4], &I.out9[5]);
    I.out9[6] = r.diag[2];
    I.out9[7] = r.diag[3];
    I.out9[8] = r.diag[4];
  }
  I.out9[0] = (r.log_ntime == last ? std::max(1.0e-8, r.diag[0]) : r.cfl) * dt;
  I.out9[1] = dt;
  I.out9[2] = I.time;
}

// Wait for all queued batch work and fold the device results into the
// instances: the first failing step and its code, as the main.F90 loop
// would have reported them, and the last logging values.
int harvest(tc_batch *b) {
  BCHK(hipStreamSynchronize(b->stream));
  for (Group &g : b->groups) {
    const int n = (int)g.ids.size();
    g.res_host.resize((size_t)n);
    BCHK(hipMemcpyAsync(g.res_host.data(), g.res, sizeof(BatchRes) * n, hipMemcpyDeviceToHost,
                        b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    for (int q = 0; q < n; q++) {
      Inst &I = b->inst[(size_t)g.ids[(size_t)q]];
      // removed since the groups were built (they are rebuilt before the
      // next step)
      if (!I.live || !I.c) continue;
      const BatchRes &r = g.res_host[(size_t)q];
      tc_ctx *c = I.c;
      if (I.done) continue;
      // earliest step with any exit condition; within that step the codes
      // are assigned in main.F90 order 165, 266/169, 333, 555 and the last wins
      int first = -1;
      auto take = [&first](int s) {
        if (s >= 0 && (first < 0 || s < first)) first = s;
      };
      take(r.fail165 - 1);
      take(r.faillog - 1);
      take(I.host_end);
      if (first < 0) continue;
      int code = 0;
      if (r.fail165 - 1 == first) code = 165;
      if (r.faillog - 1 == first) code = r.faillog_code;
      if (I.host_end == first) code = I.host_code;
      I.done = true;
      I.code = code;
      I.end_step = first;
      // a device-detected end stopped the instance after that step; the
      // host clock kept counting until this harvest
      if (first == r.fail165 - 1 || first == r.faillog - 1) I.time = r.fail_time;
      b->dirty = true;
    }
  }
  for (Group &g : b->groups) {
    for (size_t q = 0; q < g.ids.size(); q++) {
      Inst &I = b->inst[(size_t)g.ids[q]];
      if (!I.live || !I.c || q >= g.res_host.size()) continue;
      set_out9(I, g.res_host[q]);
    }
  }
  return 0;
}

StepDesc make_desc(tc_ctx *c, int ns) {
  StepDesc d;
  std::memset(&d, 0, sizeof d);
  const tc::Grid &g = c->g;
  d.g = c->gd;
  const double al = g.alm[ns], dt = c->dt;
  d.a.al = al;
  d.a.ga = g.gam[ns];
  d.a.ro = g.rom[ns];
  d.a.dt = dt;
  d.a.alre = al / c->p.ren;
  d.a.vwall_in = c->p.virotint;
  d.a.virotext = c->p.virotext;
  d.a.bc1 = 0.0;
  d.a.udxz = g.dxz * al;
  for (int q = 0; q < TC_NFIELDS; q++) d.f[q] = c->f[q];
  d.spec = c->spec;
  d.icyl = c->mask;
  d.ocyl = c->mask + c->P;
  const size_t slot = (size_t)(ns - 1) * 3;
  d.lut = c->lu + (slot + tc::LU_QT) * 4 * (size_t)c->lu_nt;
  d.lur = c->lu + (slot + tc::LU_QR) * 4 * (size_t)c->lu_nt;
  d.luz = c->lu + (slot + tc::LU_QZ) * 4 * (size_t)c->lu_nt;
  d.lu_nt = c->lu_nt;
  d.dt = dt;
  d.al = al;
  d.beta = c->beta;
  d.usdt = 1.0 / (al * dt);
  d.inv_norm = 1.0 / ((double)c->nthm * (double)c->nzm);
  d.pt = c->plan_t;
  d.pz = c->plan_z;
  d.twt = c->tw_t;
  d.twz = c->tw_z;
  d.red = c->red;
  d.red_out = c->red_out;
  // deferred correction: the same alternation of pr and pr_alt as the
  // single-instance schedule
  d.corr = (tc::defers_correction(c) && ns > 1) ? 1 : 0;
  d.corr_al = ns > 1 ? g.alm[ns - 1] : 0.0;
  d.pr_in = (d.corr && ns % 2 == 1) ? c->pr_alt : c->f[TC_PR];
  d.pr_out = d.corr ? (ns % 2 == 0 ? c->pr_alt : c->f[TC_PR]) : nullptr;
  d.limitcfl = c->p.limitcfl;
  d.resid = c->p.resid;
  d.limitvel = c->p.limitvel;
  d.virotint = c->p.virotint;
  d.virotext = c->p.virotext;
  d.del = -1.0 / (g.rc[2] - g.rc[1]);
  d.deln = -1.0 / (g.rc[c->nr] - g.rc[c->nrm]);
  d.pi = g.pi;
  return d;
}

int rebuild(tc_batch *b) {
  int rc = harvest(b);
  if (rc) return rc;
  free_groups(b);
  std::map<Key, std::vector<int>> by;
  for (size_t i = 0; i < b->inst.size(); i++) {
    Inst &I = b->inst[i];
    if (!I.live || I.done) continue;
    by[group_key(I.c)].push_back((int)i);
  }
  size_t total = 0;
  for (auto &kv : by) {
    Group g;
    g.ids = kv.second;
    g.rep = b->inst[(size_t)g.ids[0]].c;
    const int n = (int)g.ids.size();
    const int nsst = g.rep->p.nsst;
    BCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    BCHK(hipEventCreateWithFlags(&g.join, hipEventDisableTiming));
    BCHK(hipMalloc((void **)&g.dyn, sizeof(BatchDyn) * n));
    BCHK(hipMalloc((void **)&g.res, sizeof(BatchRes) * n));
    std::vector<StepDesc> h((size_t)n);
    for (int ns = 1; ns <= nsst; ns++) {
      for (int q = 0; q < n; q++) {
        tc_ctx *c = b->inst[(size_t)g.ids[(size_t)q]].c;
        h[(size_t)q] = make_desc(c, ns);
        h[(size_t)q].dyn = g.dyn + q;
        h[(size_t)q].res = g.res + q;
      }
      BCHK(hipMalloc((void **)&g.tab[ns - 1], sizeof(StepDesc) * n));
      BCHK(hipMemcpyAsync(g.tab[ns - 1], h.data(), sizeof(StepDesc) * n, hipMemcpyHostToDevice,
                          b->stream));
      BCHK(hipStreamSynchronize(b->stream));
    }
    std::vector<BatchRes> r0((size_t)n);
    for (BatchRes &r : r0) {
      std::memset(&r, 0, sizeof r);
      r.log_ntime = -1;
    }
    BCHK(hipMemcpyAsync(g.res, r0.data(), sizeof(BatchRes) * n, hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    g.dyn_off = total;
    total += (size_t)n;
    for (int q = 0; q < n; q++) {
      Inst &I = b->inst[(size_t)g.ids[(size_t)q]];
      I.group = (int)b->groups.size();
      I.slot = q;
    }
    b->groups.push_back(std::move(g));
  }
  if (total > b->dyn_cap) {
    for (int k = 0; k < tc_batch::RING; k++) {
      if (b->dyn_ev[k]) BCHK(hipEventSynchronize(b->dyn_ev[k]));
      if (b->dyn_host[k]) hipHostFree(b->dyn_host[k]);
      b->dyn_host[k] = nullptr;
      BCHK(hipHostMalloc((void **)&b->dyn_host[k], sizeof(BatchDyn) * std::max<size_t>(total, 1),
                         hipHostMallocDefault));
    }
    b->dyn_cap = total;
  }
  b->launches = 0;
  for (Group &g : b->groups) b->launches += 6 + 8 * g.rep->p.nsst;
  b->dirty = false;
  return 0;
}

// Queue one step of every group: fork the group streams from the batch
// stream, run each group's grouped kernels, join.
int enqueue_step(tc_batch *b) {
  BCHK(hipEventRecord(b->fork, b->stream));
  for (Group &g : b->groups) {
    const int n = (int)g.ids.size();
    tc_ctx *c = g.rep;
    BCHK(hipStreamWaitEvent(g.stream, b->fork, 0));
    tc::launch_cfl_check_g(c, g.tab[0], g.dyn, g.res, n, g.stream);
    for (int ns = 1; ns <= c->p.nsst; ns++) {
      const bool defer = tc::defers_correction(c);
      tc::launch_helmholtz_fused_g(c, g.tab[ns - 1], g.dyn, ns, n, g.stream, defer && ns > 1);
      tc::launch_poisson_g(c, g.tab[ns - 1], n, g.stream);
      if (!defer || ns == c->p.nsst) tc::launch_correct_both_g(c, g.tab[ns - 1], n, g.stream);
    }
    tc::launch_diagnostics_g(c, g.tab[0], g.dyn, g.res, n, g.stream);
    BCHK(hipEventRecord(g.join, g.stream));
    BCHK(hipStreamWaitEvent(b->stream, g.join, 0));
  }
  return tc::check_launch();
}

}  // namespace

extern "C" {

int tc_batch_create(tc_batch **out) {
  if (!out) return tc::set_error("null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return tc::set_error("no HIP device available");
  tc_batch *b = new tc_batch();
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  for (int k = 0; k < tc_batch::RING && e == hipSuccess; k++)
    e = hipEventCreateWithFlags(&b->dyn_ev[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    tc_batch_destroy(b);
    return tc::set_error(std::string("tc_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return 0;
}

int tc_batch_destroy(tc_batch *b) {
  if (!b) return 0;
  if (b->stream) hipStreamSynchronize(b->stream);
  free_groups(b);
  for (Inst &I : b->inst)
    if (I.live && I.c) tc_destroy(I.c);
  for (int k = 0; k < tc_batch::RING; k++) {
    if (b->dyn_ev[k]) {
      hipEventSynchronize(b->dyn_ev[k]);
      hipEventDestroy(b->dyn_ev[k]);
    }
    if (b->dyn_host[k]) hipHostFree(b->dyn_host[k]);
  }
  if (b->fork) hipEventDestroy(b->fork);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int tc_batch_add(tc_batch *b, const tc_params *p, const tc_tuning *t, int *id) {
  if (!b || !p || !id) return tc::set_error("null argument");
  if (p->idtv != 0)
    return tc::set_error("batched instances advance at a fixed dt: IDTV must be 0 (dt = DTMAX)");
  if (p->minstrt * p->multmin + p->minstrz * p->multmin != 0)
    return tc::set_error("batched instances need no-slip walls everywhere (MINSTRT*MULTMIN = MINSTRZ*MULTMIN = 0)");
  if (!(p->dtmax > 0.0)) return tc::set_error("batched instances need DTMAX > 0");
  if (p->tsta < p->tmax)
    return tc::set_error("batched instances keep no statistics: TSTA must not be below TMAX "
                         "(run statistics on a single instance, tc_step)");
  tc_ctx *c = nullptr;
  int rc = tc_create_tuned(p, t, &c);
  if (rc) return rc;
  // the batch owns time and dt: set the step-invariant tables once
  c->batch_owned = true;
  c->dt = p->dtmax;
  c->time = 0.0;
  rc = tc::ctx_begin_step(c);
  if (rc == 0 && hipStreamSynchronize(c->stream) != hipSuccess) rc = tc::set_error("tc_batch_add: stream");
  if (rc) {
    tc_destroy(c);
    return rc;
  }
  Inst I;
  I.c = c;
  I.live = true;
  b->inst.push_back(I);
  *id = (int)b->inst.size() - 1;
  b->dirty = true;
  return 0;
}

int tc_batch_remove(tc_batch *b, int id) {
  if (!b || id < 0 || id >= (int)b->inst.size() || !b->inst[(size_t)id].live)
    return tc::set_error("tc_batch_remove: no such instance");
  int rc = harvest(b);
  if (rc) return rc;
  Inst &I = b->inst[(size_t)id];
  tc_destroy(I.c);
  I.c = nullptr;
  I.live = false;
  b->dirty = true;
  return 0;
}

int tc_batch_ctx(tc_batch *b, int id, tc_ctx **ctx) {
  if (!b || !ctx || id < 0 || id >= (int)b->inst.size() || !b->inst[(size_t)id].live)
    return tc::set_error("tc_batch_ctx: no such instance");
  int rc = harvest(b);
  if (rc) return rc;
  *ctx = b->inst[(size_t)id].c;
  return 0;
}

int tc_batch_advance(tc_batch *b, int nsteps, int use_graph, int sync_every) {
  if (!b) return tc::set_error("null argument");
  if (sync_every < 1) sync_every = 1;
  // work the caller queued on the instances' own streams comes first
  for (Inst &I : b->inst)
    if (I.live && I.c) BCHK(hipStreamSynchronize(I.c->stream));
  int rc = 0;
  for (int s = 0; s < nsteps; s++) {
    if (b->dirty) {
      rc = rebuild(b);
      if (rc) return rc;
    }
    if (b->groups.empty()) break;
    // per-step scalars, computed exactly as tc_step / begin_step do
    const int k = b->ring;
    b->ring = (b->ring + 1) % tc_batch::RING;
    BCHK(hipEventSynchronize(b->dyn_ev[k]));
    BatchDyn *dh = b->dyn_host[k];
    for (Group &g : b->groups) {
      for (size_t q = 0; q < g.ids.size(); q++) {
        Inst &I = b->inst[(size_t)g.ids[q]];
        tc_ctx *c = I.c;
        const tc_params &p = c->p;
        const double pi = c->g.pi, dt = p.dtmax;
        const double vinner = p.sizepert * std::sin(2. * pi * p.omegapert * I.time);
        const double dvinner = 2. * pi * p.omegapert * p.sizepert * std::cos(2. * pi * p.omegapert * I.time) * dt;
        BatchDyn &d = dh[g.dyn_off + q];
        for (int ns = 1; ns <= 3; ns++) {
          const double al = c->g.alm[std::min(ns, p.nsst)];
          d.vwall[ns - 1] = p.virotint + vinner;
          d.bc1[ns - 1] = 0.5 * al * dt * dvinner * c->g.am1j[1] / p.ren;
        }
        const int ntime = I.steps;
        const double tnew = I.time + dt;
        d.ntime = ntime;
        d.tnew = tnew;
        d.log = (ntime == 1 || std::fmod(tnew, p.tout) < dt) ? 1 : 0;
      }
      BCHK(hipMemcpyAsync(g.dyn, dh + g.dyn_off, sizeof(BatchDyn) * g.ids.size(),
                          hipMemcpyHostToDevice, b->stream));
    }
    BCHK(hipEventRecord(b->dyn_ev[k], b->stream));
    if (use_graph) {
      if (!b->exec) {
        // relaxed: other threads (the served loop's other shape classes)
        // keep creating contexts and uploading fields while one captures
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        rc = enqueue_step(b);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc) return rc;
        BCHK(e);
        b->graph = graph;
        BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      }
      BCHK(hipGraphLaunch(b->exec, b->stream));
    } else {
      rc = enqueue_step(b);
      if (rc) return rc;
    }
    // host side of the step: time, and the ends the host can tell
    for (Group &g : b->groups)
      for (int id : g.ids) {
        Inst &I = b->inst[(size_t)id];
        const tc_params &p = I.c->p;
        const int ntime = I.steps;
        I.time = I.time + p.dtmax;
        I.steps++;
        int code = 0;
        if (I.time > p.tmax) code = 333;
        if (ntime == p.ntst) code = 555;
        if (code && I.host_end < 0) {
          I.host_end = ntime;
          I.host_code = code;
          b->dirty = true;
        }
      }
    if ((s + 1) % sync_every == 0 || s + 1 == nsteps) {
      rc = harvest(b);
      if (rc) return rc;
    }
  }
  rc = harvest(b);
  if (rc) return rc;
  int active = 0;
  for (Inst &I : b->inst)
    if (I.live && !I.done) active++;
  return active;
}

int tc_batch_status(tc_batch *b, int id, int *steps, int *code, double *out9) {
  if (!b || id < 0 || id >= (int)b->inst.size() || !b->inst[(size_t)id].live)
    return tc::set_error("tc_batch_status: no such instance");
  int rc = harvest(b);
  if (rc) return rc;
  const Inst &I = b->inst[(size_t)id];
  if (steps) *steps = I.done ? I.end_step + 1 : I.steps;
  if (code) *code = I.done ? I.code : 0;
  if (out9)
    for (int q = 0; q < 9; q++) out9[q] = I.out9[q];
  return 0;
}

// One instance's whole state at its current step, in the tc_write_restart
// format: the batch keeps each instance's step count and time itself, so
// they are handed to the context first (the CFL cache is marked stale: a
// run continued on the single-instance path evaluates it afresh).
int tc_batch_write_restart(tc_batch *b, int id, const char *path) {
  if (!b || !path || id < 0 || id >= (int)b->inst.size() || !b->inst[(size_t)id].live)
    return tc::set_error("tc_batch_write_restart: no such instance");
  int rc = harvest(b);
  if (rc) return rc;
  Inst &I = b->inst[(size_t)id];
  tc_ctx *c = I.c;
  c->time = I.time;
  c->dt = c->p.dtmax;
  c->cfl_valid = false;
  return tc_write_restart(c, path, I.done ? I.end_step + 1 : I.steps);
}

// Loads a checkpoint into an instance that has not stepped yet; it then
// continues from the checkpoint's step and time, bit for bit the run that
// wrote it.  The checkpoint's dt must be the instance's fixed dt.
int tc_batch_read_restart(tc_batch *b, int id, const char *path, int *ntime) {
  if (!b || !path || id < 0 || id >= (int)b->inst.size() || !b->inst[(size_t)id].live)
    return tc::set_error("tc_batch_read_restart: no such instance");
  int rc = harvest(b);
  if (rc) return rc;
  Inst &I = b->inst[(size_t)id];
  if (I.steps != 0 || I.done)
    return tc::set_error("tc_batch_read_restart: the instance has already stepped");
  // checked before any field is replaced, so a refused file leaves the
  // instance as it was
  double fdt = 0.0;
  int fnt = 0;
  rc = tc::restart_file_info(path, &fdt, &fnt);
  if (rc) return rc;
  if (fdt != I.c->p.dtmax)
    return tc::set_error("tc_batch_read_restart: the checkpoint's dt is not the instance's DTMAX");
  if (fnt > I.c->p.ntst)
    return tc::set_error("tc_batch_read_restart: the checkpoint is at step " + std::to_string(fnt) +
                         ", past NTST = " + std::to_string(I.c->p.ntst) + ": nothing left to run");
  int nt = 0;
  rc = tc_read_restart(I.c, path, &nt);
  if (rc) return rc;
  I.steps = nt;
  I.time = I.c->time;
  if (ntime) *ntime = nt;
  return 0;
}

int tc_batch_groups(tc_batch *b, int *ngroups, int *launches_per_step) {
  if (!b) return tc::set_error("null argument");
  if (b->dirty) {
    int rc = rebuild(b);
    if (rc) return rc;
  }
  if (ngroups) *ngroups = (int)b->groups.size();
  if (launches_per_step) *launches_per_step = b->launches;
  return 0;
}

}  // extern "C"
