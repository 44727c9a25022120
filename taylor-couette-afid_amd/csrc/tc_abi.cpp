// C-ABI of libtcamd.so: device context, state transfer and the AFiD routine
// entries.  See include/tc_abi.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <tuple>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "tc_internal.h"

#include <chrono>

namespace {

thread_local std::string g_err;

// Every reported failure also clears the runtime's last-error state: a HIP
// call that failed on the way (an allocation, a copy) must not surface again
// at the next launch check as a failed kernel launch.
int fail(const std::string &msg) {
  g_err = msg;
  (void)hipGetLastError();
  return -1;
}

// A failed call is reported here and cleared from the runtime's last-error
// state, as fail() does.
#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      g_err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
      (void)hipGetLastError();                                                           \
      return -2;                                                                         \
    }                                                                                    \
  } while (0)

#define CTXCHK(c)                                   \
  do {                                              \
    if (!(c)) return fail("null tc_ctx");           \
    if (hipSetDevice((c)->device) != hipSuccess)    \
      return fail("hipSetDevice failed");           \
  } while (0)

int launch_status() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_err = std::string("kernel launch: ") + hipGetErrorString(e);
    return -2;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Helmholtz LU tables (SolveImpEqnUpdate_QT.F90:37-63,
// SolveImpEqnUpdate_QR.F90:27-44, SolveImpEqnUpdate_QZ.F90:35-58)
// ---------------------------------------------------------------------------
void write_table(double *T, int nt, int n, std::vector<double> &dl, std::vector<double> &d,
                 std::vector<double> &du, const std::vector<double> &sc) {
  constexpr int W = tc::LU_ROW;
  std::vector<double> du2((size_t)std::max(n, 1), 0.0);
  std::vector<int> ipiv((size_t)n);
  tc::dgttrf(n, dl.data(), d.data(), du.data(), du2.data(), ipiv.data());
  std::fill(T, T + nt, 0.0);
  for (int q = 0; q < n - 1; q++) {
    T[W * q] = dl[q];
    T[W * q + 1] = du[q];
    T[W * q + 5] = (ipiv[q] != q) ? 1.0 : 0.0;
  }
  for (int q = 0; q < n - 2; q++) T[W * q + 2] = du2[q];
  for (int q = 0; q < n; q++) {
    T[W * q + 3] = 1.0 / d[q];
    T[W * q + 4] = sc[q];
    T[W * q + 6] = d[q];
  }
}

// Row-normalised matrix of the q_theta / q_z solves (mask variant ic, oc).
void build_wall_table(const tc::Grid &g, int comp, int ic, int oc, double betadx, double *T, int nt) {
  const int n = g.nrm;
  const bool is_qt = (comp == tc::LU_QT);
  const std::vector<double> &ap = is_qt ? g.ap1j : g.ap3j;
  const std::vector<double> &am = is_qt ? g.am1j : g.am3j;
  const std::vector<double> &ac = is_qt ? g.ac1j : g.ac3j;
  const double ac_fs1 = is_qt ? g.ac1j_fs[1] : g.ac3j_fs[1];
  const double ac_fs2 = is_qt ? g.ac1j_fs[2] : g.ac3j_fs[2];
  std::vector<double> amkl((size_t)n + 2), ackl((size_t)n + 2), apkl((size_t)n + 2);
  const double fic = (double)ic, foc = (double)oc;
  double actot = (ac[1] * fic + ac_fs1 * (1 - fic));
  double ackl_b = 1.0 / (1.0 - actot * betadx);
  ackl[1] = 1.0;
  apkl[1] = -ap[1] * betadx * ackl_b;
  for (int kc = 2; kc <= n - 1; kc++) {
    ackl_b = 1.0 / (1.0 - ac[kc] * betadx);
    amkl[kc] = -am[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -ap[kc] * betadx * ackl_b;
  }
  actot = (ac[n] * foc + ac_fs2 * (1 - foc));
  ackl_b = 1.0 / (1.0 - actot * betadx);
  amkl[n] = -am[n] * betadx * ackl_b;
  ackl[n] = 1.0;
  apkl[n] = 0.0;
  std::vector<double> dl((size_t)n), d((size_t)n), du((size_t)n), sc((size_t)n);
  for (int q = 0; q < n - 1; q++) {
    dl[q] = amkl[q + 2];
    du[q] = apkl[q + 1];
  }
  for (int q = 0; q < n; q++) d[q] = ackl[q + 1];
  // right-hand side row scaling uses the unblended diagonal
  for (int kc = 1; kc <= n; kc++) sc[kc - 1] = 1.0 / (1. - ac[kc] * betadx);
  write_table(T, nt, n, dl, d, du, sc);
}

void build_qr_table(const tc::Grid &g, double betadx, double *T, int nt) {
  const int n = g.nr, nrm = g.nrm;
  std::vector<double> amkl((size_t)n + 2), ackl((size_t)n + 2), apkl((size_t)n + 2);
  amkl[1] = 0.;
  apkl[1] = 0.;
  ackl[1] = 1.;
  for (int kc = 2; kc <= nrm; kc++) {
    const double ackl_b = 1.0 / (1.0 - g.ac2j[kc] * betadx);
    amkl[kc] = -g.am2j[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -g.ap2j[kc] * betadx * ackl_b;
  }
  amkl[n] = 0.;
  apkl[n] = 0.;
  ackl[n] = 1.;
  std::vector<double> dl((size_t)n), d((size_t)n), du((size_t)n), sc((size_t)n, 0.0);
  for (int q = 0; q < n - 1; q++) {
    dl[q] = amkl[q + 2];
    du[q] = apkl[q + 1];
  }
  for (int q = 0; q < n; q++) d[q] = ackl[q + 1];
  for (int kc = 2; kc <= nrm; kc++) sc[kc - 1] = 1.0 / (1.0 - g.ac2j[kc] * betadx);
  write_table(T, nt, n, dl, d, du, sc);
}

int refresh_tables(tc_ctx *c) {
  const int nsst = c->p.nsst;
  // with every wall mask 1 only the no-slip variant (3) is ever read
  const bool all_variants = !c->masks_uniform_one;
  bool any = !c->lu_valid || (all_variants && !c->lu_all_variants);
  for (int s = 0; s < nsst; s++)
    if (c->lu_key[s] != c->beta * c->g.alm[s + 1]) any = true;
  if (!any) return 0;
  // the previous upload must have left the staging buffer
  HIPCHK(hipEventSynchronize(c->lu_copied));
  const size_t per_slot = (size_t)3 * 4 * c->lu_nt;
  for (int s = 0; s < nsst; s++) {
    const double betadx = c->beta * c->g.alm[s + 1];
    double *base = c->lu_host + (size_t)s * per_slot;
    for (int v = all_variants ? 0 : 3; v < 4; v++) {
      build_wall_table(c->g, tc::LU_QT, v & 1, v >> 1, betadx,
                       base + ((size_t)tc::LU_QT * 4 + v) * c->lu_nt, c->lu_nt);
      build_wall_table(c->g, tc::LU_QZ, v & 1, v >> 1, betadx,
                       base + ((size_t)tc::LU_QZ * 4 + v) * c->lu_nt, c->lu_nt);
    }
    build_qr_table(c->g, betadx, base + (size_t)tc::LU_QR * 4 * c->lu_nt, c->lu_nt);
    c->lu_key[s] = betadx;
  }
  HIPCHK(hipMemcpyAsync(c->lu, c->lu_host, sizeof(double) * per_slot * nsst, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipEventRecord(c->lu_copied, c->stream));
  c->lu_valid = true;
  c->lu_all_variants = all_variants;
  return 0;
}

int upload_masks(tc_ctx *c) {
  std::vector<unsigned char> m((size_t)2 * c->P);
  bool ones = true;
  for (int i = 0; i < c->nthm; i++)
    for (int j = 0; j < c->nzm; j++) {
      const size_t f = (size_t)j + (size_t)c->nzm * i;
      const size_t d = (size_t)j * c->nthm + i;
      m[d] = (unsigned char)(c->icyl[f] != 0);
      m[c->P + d] = (unsigned char)(c->ocyl[f] != 0);
      ones = ones && c->icyl[f] == 1 && c->ocyl[f] == 1;
    }
  c->masks_uniform_one = ones;
  HIPCHK(hipMemcpyAsync(c->mask, m.data(), m.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// SlipBCs at the current time; re-uploads only when the pattern can change
int apply_slip_bcs(tc_ctx *c) {
  if (c->p.minstrt * c->p.multmin + c->p.minstrz * c->p.multmin == 0 && c->masks_uniform_one)
    return 0;
  tc::slip_bcs(c->p, c->g, c->time, c->icyl, c->ocyl);
  return upload_masks(c);
}

int dev_alloc(tc_ctx *c, void **p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes));
  c->device_bytes += bytes;
  return 0;
}

void destroy_ctx(tc_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (int q = 0; q < TC_NFIELDS; q++)
    if (c->f[q]) hipFree(c->f[q]);
  if (c->pr_alt) hipFree(c->pr_alt);
  if (c->spec) hipFree(c->spec);
  if (c->gvbuf) hipFree(c->gvbuf);
  if (c->mask) hipFree(c->mask);
  if (c->lu) hipFree(c->lu);
  if (c->lu_host) hipHostFree(c->lu_host);
  if (c->lu_copied) hipEventDestroy(c->lu_copied);
  if (c->tw_t) hipFree(c->tw_t);
  if (c->tw_z) hipFree(c->tw_z);
  if (c->red) hipFree(c->red);
  if (c->red_out) hipFree(c->red_out);
  if (c->stats) hipFree(c->stats);
  if (c->stat_part) hipFree(c->stat_part);
  if (c->dump) hipFree(c->dump);
  if (c->red_host) hipHostFree(c->red_host);
  for (auto &t : c->timers) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  for (auto e : c->event_pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int read_reduction(tc_ctx *c, int nv, double *out) {
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nv; q++) out[q] = c->red_host[q];
  return 0;
}

// Deferred correction (fused RK3 with an even nthm): substep 2 reads the
// uncorrected pressure from pr and writes the corrected one to pr_alt,
// substep 3 the other way round, and the last substep's correction is a
// pass of its own, so every step ends with its fields where the per-substep
// schedule leaves them.  The two schedules agree to rounding: the corrected
// values are the same bits, but the compiler forms the Helmholtz kernel's
// multiply-adds differently in its two variants.
int run_substep(tc_ctx *c) {
  const bool defer = c->fused && ::tc::defers_correction(c);
  if (c->fused && defer && c->ns > 1) {
    const bool even = (c->ns % 2 == 0);
    ::tc::launch_helmholtz_fused_corr(c, even ? c->f[TC_PR] : c->pr_alt,
                                    even ? c->pr_alt : c->f[TC_PR], c->corr_al);
  } else if (c->fused) {
    tc::launch_helmholtz_fused(c);
  } else {
    tc::launch_explicit_qt(c);
    tc::launch_explicit_qr(c);
    tc::launch_explicit_qz(c);
    tc::launch_implicit_qt(c);
    tc::launch_implicit_qr(c);
    tc::launch_implicit_qz(c);
  }
  if (c->fused) {
    tc::launch_poisson(c, true, 1.0 / (c->al * c->dt));
  } else {
    tc::launch_divergence(c, 1.0 / (c->al * c->dt));
    tc::launch_poisson(c);
  }
  if (c->fused && defer && c->ns < c->p.nsst) {
    c->corr_al = c->al;  // applied by the next substep's Helmholtz kernel
  } else if (c->fused && c->diag_in_correction && c->ns == c->p.nsst) {
    tc::launch_correct_diag(c);
  } else if (c->fused) {
    tc::launch_correct_both(c);
  } else {
    tc::launch_correct_velocity(c, c->al);
    tc::launch_correct_pressure(c);
  }
  return launch_status();
}

int begin_step(tc_ctx *c) {
  // TimeMarcher.F90:20-23
  c->beta = c->dt / c->p.ren * 0.5;
  const double pi = c->g.pi;
  c->vinner_t = c->p.sizepert * std::sin(2. * pi * c->p.omegapert * c->time);
  c->dvinner_t = 2. * pi * c->p.omegapert * c->p.sizepert * std::cos(2. * pi * c->p.omegapert * c->time) * c->dt;
  int rc = apply_slip_bcs(c);
  if (rc) return rc;
  return refresh_tables(c);
}

void set_substep(tc_ctx *c, int ns) {
  c->ns = ns;
  c->al = c->g.alm[ns];
  c->ga = c->g.gam[ns];
  c->ro = c->g.rom[ns];
}


// Effective tuning: explicit nonzero fields first, then the TC_* environment
// overrides, then the defaults (which come from sweeps at 512x512x385).
// nrm: the radial size, for the defaults that depend on it
tc_tuning resolve_tuning(const tc_tuning *in, int nrm) {
  tc_tuning t{};
  if (in) t = *in;
  auto env = [](const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
  };
  if (t.helm_lpt == 0) t.helm_lpt = std::max(0, std::min(2, env("TC_HELM_LPT", 0)));
  if (t.fft_t_lds == 0) t.fft_t_lds = env("TC_FFT_T_LDS", 32768);
  if (t.fft_z_lds == 0) t.fft_z_lds = env("TC_FFT_Z_LDS", 65536);
  if (t.fft_z_nt == 0) t.fft_z_nt = env("TC_FFT_Z_NT", 512) == 256 ? 256 : 512;
  if (t.fft_pad == 0) {
    const int v = env("TC_FFT_PAD", 1);
    t.fft_pad = (v == 0) ? -1 : std::max(1, std::min(3, v));
  }
  // rows per thread of the radial Poisson solve: 13 up to nrm = 512, 26
  // beyond, where the 2*PT-row reduced system of 13-row partitions grows
  // long (at 1024x1024x769: 4.0 ms per launch with 26 against 7.3 with 13;
  // at 512x512x385 13 is the faster)
  if (t.poisson_pq == 0) t.poisson_pq = env("TC_POISSON_PQ", nrm > 512 ? 26 : 13) == 26 ? 26 : 13;
  if (t.defer == 0) {
    const int v = env("TC_DEFER_CORRECTION", 1);
    t.defer = v == 0 ? -1 : (v == 2 ? 2 : 1);
  }
  return t;
}

// theta: row-pair FFTs within fft_t_lds bytes of LDS per block; z: column
// tiles within fft_z_lds, as wide as that allows (wider tiles = longer row
// segments)
std::string apply_plans(tc_ctx *c, const tc_tuning &t) {
  if (t.poisson_pq != 13 && t.poisson_pq != 26)
    return "poisson_pq must be 13 or 26 (got " + std::to_string(t.poisson_pq) + ")";
  if (t.helm_lpt < 0 || t.helm_lpt > 2)
    return "helm_lpt must be 0, 1 or 2 (got " + std::to_string(t.helm_lpt) + ")";
  if (t.defer != -1 && t.defer != 1 && t.defer != 2)
    return "defer must be 0, 1, 2 or -1 (got " + std::to_string(t.defer) + ")";
  tc::FFTPlan pt, pz;
  std::string e = tc::make_fft_plan(c->p.nthm, pt, t.fft_t_lds, 256);
  if (e.empty()) e = tc::make_fft_plan(c->p.nzm, pz, t.fft_z_lds, t.fft_z_nt == 256 ? 256 : 512);
  if (!e.empty()) return e;
  // one LDS padding slot per 8 (1), 4 (2) or 16 (3) points; -1: none
  pt.padsh = pz.padsh = t.fft_pad < 0 ? 31 : (t.fft_pad == 2 ? 2 : (t.fft_pad == 3 ? 4 : 3));
  c->plan_t = pt;
  c->plan_z = pz;
  c->tuning = t;
  return std::string();
}

// CalcTorque.F90:26-55: normalisation of the masked wall sums
void torque_from_sums(const tc_ctx *c, double slow, double sup, double *inner, double *outer) {
  const tc_params &p = c->p;
  const double reninv = 1.0 / p.ren;
  const double jlam = 2.0 * p.rint * p.rint * p.rext * p.rext * (p.virotint / p.rint - p.virotext / p.rext) /
                      (p.rext * p.rext - p.rint * p.rint) * reninv;
  const double jlami = 1.0 / jlam;
  const double rc2 = c->g.rc[2], rcn = c->g.rc[c->nrm];
  const double lo = slow * (rc2 * rc2 * rc2) * reninv * jlami;
  const double up = sup * (rcn * rcn * rcn) * reninv * jlami;
  const double my_surface = 1.0 / (c->nthm * c->nzm);
  *inner = lo * my_surface;
  *outer = up * my_surface;
}

}  // namespace

namespace tc {
// tc_tuning::defer: 1 defers from 8 Mi grid points up, where the saved
// passes pay; below that a lone grid lives in the caches and the
// correction's extra instructions cost more than they save (a batch of many
// small instances is another matter).  2: always; -1: never.  The spare
// pressure array is allocated where the deferral applies.
int set_deferral(tc_ctx *c, int mode) {
  c->defer_corr = mode == 2 || (mode == 1 && (long long)c->nthm * c->nzm * c->nrm >= (8LL << 20));
  if (c->defer_corr && c->p.nsst == 3 && c->nthm % 2 == 0 && !c->pr_alt) {
    int rc = dev_alloc(c, (void **)&c->pr_alt, sizeof(double) * (size_t)c->N);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->pr_alt, 0, sizeof(double) * (size_t)c->N, c->stream));
  }
  return 0;
}

bool defers_correction(const tc_ctx *c) {
  return c->defer_corr && c->pr_alt && c->p.nsst == 3 && c->nthm % 2 == 0;
}
}  // namespace tc

namespace tc {

int ctx_begin_step(tc_ctx *c) { return begin_step(c); }
void ctx_set_substep(tc_ctx *c, int ns) { set_substep(c, ns); }
void ctx_torque(const tc_ctx *c, double slow, double sup, double *inner, double *outer) {
  torque_from_sums(c, slow, sup, inner, outer);
}
int set_error(const std::string &msg) { return fail(msg); }
int check_launch() { return launch_status(); }

void timer_begin(tc_ctx *c, const char *name) {
  c->timer_open = false;
  if (!c->timing) return;
  if (!c->timing_only.empty() && c->timing_only != name) return;
  c->timer_open = true;
  Timer t;
  t.name = name;
  for (int q = 0; q < 2; q++) {
    hipEvent_t e;
    if (!c->event_pool.empty()) {
      e = c->event_pool.back();
      c->event_pool.pop_back();
    } else {
      hipEventCreate(&e);
    }
    (q == 0 ? t.a : t.b) = e;
  }
  hipEventRecord(t.a, c->stream);
  c->timers.push_back(t);
}

void timer_end(tc_ctx *c) {
  if (!c->timer_open || c->timers.empty()) return;
  hipEventRecord(c->timers.back().b, c->stream);
  c->timer_open = false;
}

}  // namespace tc

extern "C" {

const char *tc_last_error(void) { return g_err.c_str(); }

int tc_grid_vector(const tc_params *p, int id, double *out, int n) {
  if (!p || !out) return fail("null argument");
  tc::Grid g;
  std::string e = tc::build_grid(*p, g);
  if (!e.empty()) return fail(e);
  int m = tc::grid_vector(g, id, out, n);
  if (m < 0) return fail("unknown grid vector id");
  return m;
}

int tc_create(const tc_params *p, tc_ctx **out) { return tc_create_tuned(p, nullptr, out); }

int tc_create_tuned(const tc_params *p, const tc_tuning *tun, tc_ctx **out) {
  if (!p || !out) return fail("null argument");
  *out = nullptr;
  tc_ctx *c = new (std::nothrow) tc_ctx();
  if (!c) return fail("out of host memory");
  c->p = *p;
  std::string e = tc::build_grid(*p, c->g);
  const tc_tuning t = resolve_tuning(tun, p->nrm);
  if (e.empty()) e = apply_plans(c, t);
  if (!e.empty()) {
    delete c;
    return fail(e);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete c;
    return fail("no HIP device available");
  }
  hipGetDevice(&c->device);
  c->nthm = p->nthm;
  c->nzm = p->nzm;
  c->nrm = p->nrm;
  c->nr = p->nrm + 1;
  c->nth2 = p->nthm / 2 + 1;
  c->sstr = (c->nth2 + 7) / 8 * 8;  // spectral rows start on 128-byte lines
  c->P = (long long)c->nzm * c->nthm;
  c->N = (long long)c->nr * c->P;
  int rc = 0;
#define TRY(x)              \
  do {                      \
    rc = (x);               \
    if (rc) {               \
      destroy_ctx(c);       \
      return rc;            \
    }                       \
  } while (0)
#define HTRY(x)                                                          \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      g_err = std::string(#x) + ": " + hipGetErrorString(e_);            \
      (void)hipGetLastError();                                           \
      destroy_ctx(c);                                                    \
      return -2;                                                         \
    }                                                                    \
  } while (0)
  HTRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int q = 0; q < TC_NFIELDS; q++) {
    TRY(dev_alloc(c, (void **)&c->f[q], sizeof(double) * (size_t)c->N));
    HTRY(hipMemsetAsync(c->f[q], 0, sizeof(double) * (size_t)c->N, c->stream));
  }
  TRY(tc::set_deferral(c, t.defer));
  const size_t M = (size_t)c->nzm * c->sstr;
  TRY(dev_alloc(c, (void **)&c->spec, sizeof(double2) * M * c->nrm));
  // grid vectors
  {
    const tc::Grid &g = c->g;
    const std::vector<double> *vs[] = {&g.rc,   &g.rm,    &g.usrc,  &g.usrm,  &g.udx2rm,
                                       &g.udx2m, &g.udx2c, &g.rudx2c, &g.udx1vis1, &g.udx1vis2,
                                       &g.udx1qm, &g.pvol, &g.ap1j,  &g.am1j,  &g.ac1j,
                                       &g.ap2je, &g.am2je, &g.ac2je, &g.ap3j,  &g.am3j,
                                       &g.ac3j,  &g.apph,  &g.amph,  &g.acph,  &g.ak1, &g.ak3,
                                       &g.udx2mm, &g.udx2cc};
    const int nv = (int)(sizeof vs / sizeof vs[0]);
    size_t stride = (size_t)std::max(c->nr, std::max(c->nthm, c->nzm)) + 2;
    stride = (stride + 7) / 8 * 8;  // keeps the plane records 64-byte aligned
    constexpr size_t HPW = sizeof(tc::HelmPlane) / sizeof(double);
    std::vector<double> h(stride * nv + HPW * (size_t)(c->nr + 1), 0.0);
    for (int q = 0; q < nv; q++) std::copy(vs[q]->begin(), vs[q]->end(), h.begin() + q * stride);
    tc::HelmPlane *hp = reinterpret_cast<tc::HelmPlane *>(h.data() + stride * nv);
    for (int k = 1; k <= c->nrm; k++) {
      tc::HelmPlane &e = hp[k];
      e.udxt_m = g.dxt * 0.25 * g.usrm[k];
      e.udxtq_m = g.dxtq / (p->ren * (g.rm[k] * g.rm[k]));
      e.udx2rm = g.udx2rm[k];
      e.usrm = g.usrm[k];
      e.usrc = g.usrc[k];
      e.usrc_p = g.usrc[k + 1];
      e.udx1vis1 = g.udx1vis1[k];
      e.udxt_c = g.dxt * 0.25 * g.usrc[k];
      e.udxtq_c = g.dxtq / (p->ren * (g.rc[k] * g.rc[k]));
      e.usrc_m = g.usrc[k - 1];
      e.udx2c = g.udx2c[k];
      e.udx1vis2 = g.udx1vis2[k];
      e.rc = g.rc[k];
      e.ap1j = g.ap1j[k];
      e.ac1j = g.ac1j[k];
      e.am1j = g.am1j[k];
      e.ap2je = g.ap2je[k];
      e.ac2je = g.ac2je[k];
      e.am2je = g.am2je[k];
      e.ap3j = g.ap3j[k];
      e.ac3j = g.ac3j[k];
      e.am3j = g.am3j[k];
      e.udxzq = g.dxzq / p->ren;
    }
    TRY(dev_alloc(c, (void **)&c->gvbuf, sizeof(double) * h.size()));
    HTRY(hipMemcpyAsync(c->gvbuf, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                        c->stream));
    HTRY(hipStreamSynchronize(c->stream));
    const double **dst[] = {&c->gd.rc,   &c->gd.rm,    &c->gd.usrc,  &c->gd.usrm,  &c->gd.udx2rm,
                            &c->gd.udx2m, &c->gd.udx2c, &c->gd.rudx2c, &c->gd.udx1vis1,
                            &c->gd.udx1vis2, &c->gd.udx1qm, &c->gd.pvol, &c->gd.ap1j,
                            &c->gd.am1j, &c->gd.ac1j, &c->gd.ap2je, &c->gd.am2je, &c->gd.ac2je,
                            &c->gd.ap3j, &c->gd.am3j, &c->gd.ac3j, &c->gd.apph, &c->gd.amph,
                            &c->gd.acph, &c->gd.ak1, &c->gd.ak3, &c->gd.udx2mm, &c->gd.udx2cc};
    for (int q = 0; q < nv; q++) *dst[q] = c->gvbuf + q * stride;
    c->gd.hp = reinterpret_cast<const tc::HelmPlane *>(c->gvbuf + stride * nv);
    c->gd.dxt = g.dxt;
    c->gd.dxz = g.dxz;
    c->gd.dxtq = g.dxtq;
    c->gd.dxzq = g.dxzq;
    c->gd.ren = p->ren;
    c->gd.romeg = p->romeg;
    c->gd.ap1j_fs2 = g.ap1j_fs[2];
    c->gd.ac1j_fs1 = g.ac1j_fs[1];
    c->gd.ac1j_fs2 = g.ac1j_fs[2];
    c->gd.am1j_fs1 = g.am1j_fs[1];
    c->gd.ap3j_fs2 = g.ap3j_fs[2];
    c->gd.ac3j_fs1 = g.ac3j_fs[1];
    c->gd.ac3j_fs2 = g.ac3j_fs[2];
    c->gd.am3j_fs1 = g.am3j_fs[1];
    c->gd.nthm = c->nthm;
    c->gd.nzm = c->nzm;
    c->gd.nrm = c->nrm;
    c->gd.nr = c->nr;
    c->gd.nth2 = c->nth2;
    c->gd.sstr = c->sstr;
    c->gd.P = c->P;
  }
  // masks (SlipBCs at time 0)
  TRY(dev_alloc(c, (void **)&c->mask, (size_t)2 * c->P));
  tc::slip_bcs(c->p, c->g, 0.0, c->icyl, c->ocyl);
  TRY(upload_masks(c));
  // LU tables
  c->lu_nt = tc::LU_ROW * c->nr;
  {
    const size_t n = (size_t)3 * 3 * 4 * c->lu_nt;
    TRY(dev_alloc(c, (void **)&c->lu, sizeof(double) * n));
    HTRY(hipHostMalloc((void **)&c->lu_host, sizeof(double) * n, hipHostMallocDefault));
    HTRY(hipEventCreateWithFlags(&c->lu_copied, hipEventDisableTiming));
    HTRY(hipEventRecord(c->lu_copied, c->stream));
  }
  // twiddles
  {
    std::vector<double> tw;
    tc::fill_twiddles(c->nthm, tw);
    TRY(dev_alloc(c, (void **)&c->tw_t, sizeof(double) * tw.size()));
    HTRY(hipMemcpyAsync(c->tw_t, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice,
                        c->stream));
    HTRY(hipStreamSynchronize(c->stream));
    tc::fill_twiddles(c->nzm, tw);
    TRY(dev_alloc(c, (void **)&c->tw_z, sizeof(double) * tw.size()));
    HTRY(hipMemcpyAsync(c->tw_z, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice,
                        c->stream));
    HTRY(hipStreamSynchronize(c->stream));
  }
  // reductions
  // per-block partials of the reductions: one block of 256 per (j, i) column group
  c->red_cap = 40 * (size_t)((c->P + 255) / 256);
  TRY(dev_alloc(c, (void **)&c->red, sizeof(double) * c->red_cap));
  TRY(dev_alloc(c, (void **)&c->red_out, sizeof(double) * 16));
  HTRY(hipHostMalloc((void **)&c->red_host, sizeof(double) * 16, hipHostMallocDefault));
  // statistics accumulators (StatRoutines.F90 InitStats: zero)
  {
    const size_t m = (size_t)c->nrm * c->nzm;
    TRY(dev_alloc(c, (void **)&c->stats, sizeof(double) * TC_NSTATS * m));
    TRY(dev_alloc(c, (void **)&c->stat_part, sizeof(double) * m));
    HTRY(hipMemsetAsync(c->stats, 0, sizeof(double) * TC_NSTATS * m, c->stream));
    c->last_nute = std::nan("");
  }
  set_substep(c, 1);
  c->helm_lpt = t.helm_lpt ? t.helm_lpt : tc::choose_helm_lpt(c);
  c->poisson_pq = t.poisson_pq;
  {
    const char *dfs = std::getenv("TC_DIVFFT_STAGE");
    c->divfft_stage = dfs ? std::atoi(dfs) : 1;
  }
  TRY(refresh_tables(c));
  HTRY(hipStreamSynchronize(c->stream));
#undef TRY
#undef HTRY
  *out = c;
  return 0;
}

int tc_set_tuning(tc_ctx *c, const tc_tuning *t) {
  CTXCHK(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  const tc_tuning r = resolve_tuning(t, c->nrm);
  const std::string e = apply_plans(c, r);
  if (!e.empty()) return fail(e);
  c->helm_lpt = r.helm_lpt ? r.helm_lpt : tc::choose_helm_lpt(c);
  c->poisson_pq = r.poisson_pq;
  return tc::set_deferral(c, r.defer);
}

int tc_get_tuning(tc_ctx *c, tc_tuning *t) {
  CTXCHK(c);
  if (!t) return fail("null argument");
  *t = c->tuning;
  t->helm_lpt = c->helm_lpt;
  t->poisson_pq = c->poisson_pq;
  return 0;
}

// Coordinate descent over the schedule knobs on one context of the shape:
// each candidate is timed over `steps` fused RK3 steps (after one untimed
// step) and kept if it is at least 2% faster.  The fields stay zero, which
// exercises the same instruction stream as any other state.
int tc_autotune(const tc_params *p, int steps, tc_tuning *best) {
  if (!p || !best) return fail("null argument");
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, tc_tuning> cache;
  const auto key = std::make_tuple(p->nthm, p->nzm, p->nrm);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *best = it->second;
      return 0;
    }
  }
  if (steps < 1) steps = 1;
  tc_params q = *p;
  q.idtv = 0;
  if (!(q.dtmax > 0.0)) q.dtmax = 1e-4;
  tc_ctx *c = nullptr;
  int rc = tc_create(&q, &c);
  if (rc) return rc;
  c->dt = q.dtmax;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    tc_destroy(c);
    return fail("tc_autotune: event");
  }
  auto run = [&](const tc_tuning &t, double &ms) -> int {
    int r = tc_set_tuning(c, &t);
    if (r) return r;
    r = tc_time_marcher(c);
    if (r) return r;
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int s = 0; s < steps; s++) {
      r = tc_time_marcher(c);
      if (r) return r;
    }
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, e0, e1));
    ms = f;
    return 0;
  };
  tc_tuning cur{};
  tc_get_tuning(c, &cur);
  double best_ms = 0.0;
  rc = run(cur, best_ms);
  struct Knob {
    int tc_tuning::*field;
    std::vector<int> values;
  };
  const Knob knobs[] = {{&tc_tuning::helm_lpt, {1, 2}},
                        {&tc_tuning::poisson_pq, {26, 13}},
                        {&tc_tuning::fft_t_lds, {16384, 32768, 65536}},
                        {&tc_tuning::fft_z_lds, {32768, 65536}},
                        {&tc_tuning::fft_z_nt, {256, 512}},
                        {&tc_tuning::fft_pad, {1, -1}},
                        {&tc_tuning::defer, {2, -1}}};
  for (const Knob &k : knobs) {
    if (rc) break;
    for (int v : k.values) {
      if (cur.*k.field == v) continue;
      tc_tuning cand = cur;
      cand.*k.field = v;
      double ms = 0.0;
      if (run(cand, ms) != 0) {
        hipGetLastError();  // a schedule that does not fit is skipped
        continue;
      }
      if (ms < 0.98 * best_ms) {
        best_ms = ms;
        cur = cand;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  tc_destroy(c);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = cur;
  }
  *best = cur;
  return 0;
}

namespace {
struct RestartHeader {
  char magic[8];
  int32_t nthm, nzm, nrm, nsst, ntime, cfl_valid, istr, pad;
  double time, dt, cfl, rint, rext, alx3, lamb, str;
};
static_assert(sizeof(RestartHeader) == 104, "restart header layout");
const int kRestartFields[] = {TC_QT, TC_QR, TC_QZ, TC_PR, TC_RU1, TC_RU2, TC_RU3};
}  // namespace

extern "C++" {
namespace tc {
// The dt and step a restart file was written with, from its header alone
int restart_file_info(const char *path, double *dt, int *ntime) {
  std::FILE *f = path ? std::fopen(path, "rb") : nullptr;
  if (!f) return fail(std::string("cannot open ") + (path ? path : "(null)"));
  RestartHeader h;
  const bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, "TCAFIDR1", 8) == 0;
  std::fclose(f);
  if (!ok) return fail(std::string(path) + " is not a restart file of this library");
  *dt = h.dt;
  *ntime = h.ntime;
  return 0;
}
}  // namespace tc
}

int tc_write_restart(tc_ctx *c, const char *path, int ntime) {
  CTXCHK(c);
  if (!path) return fail("null path");
  RestartHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, "TCAFIDR1", 8);
  h.nthm = c->nthm;
  h.nzm = c->nzm;
  h.nrm = c->nrm;
  h.nsst = c->p.nsst;
  h.ntime = ntime;
  h.cfl_valid = c->cfl_valid ? 1 : 0;
  h.istr = c->p.istr;
  h.time = c->time;
  h.dt = c->dt;
  h.cfl = c->cfl_cached;
  h.rint = c->p.rint;
  h.rext = c->p.rext;
  h.alx3 = c->p.alx3;
  h.lamb = c->p.lamb;
  h.str = c->p.str;
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  std::vector<double> buf((size_t)c->N);
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q : kRestartFields) {
    if (!ok) break;
    HIPCHK(hipMemcpyAsync(buf.data(), c->f[q], sizeof(double) * buf.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ok = std::fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail(std::string("short write to ") + path);
}

int tc_read_restart(tc_ctx *c, const char *path, int *ntime) {
  CTXCHK(c);
  if (!path) return fail("null path");
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path);
  RestartHeader h;
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, "TCAFIDR1", 8) != 0) {
    std::fclose(f);
    return fail(std::string(path) + " is not a restart file of this library");
  }
  if (h.nthm != c->nthm || h.nzm != c->nzm || h.nrm != c->nrm || h.nsst != c->p.nsst ||
      h.istr != c->p.istr || h.rint != c->p.rint || h.rext != c->p.rext || h.alx3 != c->p.alx3 ||
      h.lamb != c->p.lamb || h.str != c->p.str) {
    std::fclose(f);
    return fail(std::string(path) + " was written for another grid or geometry");
  }
  std::vector<double> buf((size_t)c->N);
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q : kRestartFields) {
    if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
      std::fclose(f);
      return fail(std::string(path) + " is truncated");
    }
    HIPCHK(hipMemcpyAsync(c->f[q], buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  std::fclose(f);
  c->time = h.time;
  c->dt = h.dt;
  c->cfl_valid = h.cfl_valid != 0;
  c->cfl_cached = h.cfl;
  if (ntime) *ntime = h.ntime;
  return 0;
}

int tc_regrid_field(tc_ctx *c, int field, const double *old, int ntho, int nzo, int nro, int istro,
                    double stro) {
  CTXCHK(c);
  return tc::regrid_field(c, field, old, ntho, nzo, nro, istro, stro);
}

int tc_destroy(tc_ctx *c) {
  destroy_ctx(c);
  return 0;
}

static int check_field(tc_ctx *c, int field) {
  if (field < 0 || field >= TC_NFIELDS) return fail("field id out of range");
  return 0;
}

int tc_upload_field(tc_ctx *c, int field, const double *host, int layout) {
  CTXCHK(c);
  if (check_field(c, field)) return -1;
  if (!host) return fail("null host pointer");
  c->cfl_valid = false;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (layout == TC_LAYOUT_DEVICE) {
    HIPCHK(hipMemcpyAsync(c->f[field], host, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice,
                           c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  if (layout != TC_LAYOUT_FORTRAN) return fail("unknown layout");
  std::vector<double> d((size_t)c->N);
  const int nr = c->nr, nzm = c->nzm, nthm = c->nthm;
  for (int i = 0; i < nthm; i++)
    for (int j = 0; j < nzm; j++) {
      const double *src = host + ((size_t)j + (size_t)nzm * i) * nr;
      for (int k = 0; k < nr; k++) d[(size_t)k * c->P + (size_t)j * nthm + i] = src[k];
    }
  HIPCHK(hipMemcpyAsync(c->f[field], d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice,
                         c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int tc_download_field(tc_ctx *c, int field, double *host, int layout) {
  CTXCHK(c);
  if (check_field(c, field)) return -1;
  if (!host) return fail("null host pointer");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (layout == TC_LAYOUT_DEVICE) {
    HIPCHK(hipMemcpyAsync(host, c->f[field], sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost,
                           c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  if (layout != TC_LAYOUT_FORTRAN) return fail("unknown layout");
  std::vector<double> d((size_t)c->N);
  HIPCHK(hipMemcpyAsync(d.data(), c->f[field], sizeof(double) * d.size(), hipMemcpyDeviceToHost,
                         c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nr = c->nr, nzm = c->nzm, nthm = c->nthm;
  for (int i = 0; i < nthm; i++)
    for (int j = 0; j < nzm; j++) {
      double *dst = host + ((size_t)j + (size_t)nzm * i) * nr;
      for (int k = 0; k < nr; k++) dst[k] = d[(size_t)k * c->P + (size_t)j * nthm + i];
    }
  return 0;
}

int tc_set_masks(tc_ctx *c, const int *icylsl, const int *ocylsl) {
  CTXCHK(c);
  if (!icylsl || !ocylsl) return fail("null mask pointer");
  const size_t n = (size_t)c->P;
  for (size_t q = 0; q < n; q++)
    if ((icylsl[q] != 0 && icylsl[q] != 1) || (ocylsl[q] != 0 && ocylsl[q] != 1))
      return fail("masks must hold 0 or 1");
  c->icyl.assign(icylsl, icylsl + n);
  c->ocyl.assign(ocylsl, ocylsl + n);
  const int rc = upload_masks(c);
  if (rc) return rc;
  // slip walls need the other three LU variants of the current tables
  return c->lu_valid ? refresh_tables(c) : 0;
}

int tc_get_masks(tc_ctx *c, int *icylsl, int *ocylsl) {
  CTXCHK(c);
  std::copy(c->icyl.begin(), c->icyl.end(), icylsl);
  std::copy(c->ocyl.begin(), c->ocyl.end(), ocylsl);
  return 0;
}

int tc_set_time(tc_ctx *c, double time, double dt) {
  CTXCHK(c);
  c->time = time;
  c->dt = dt;
  return 0;
}

int tc_get_scalars(tc_ctx *c, double *o) {
  CTXCHK(c);
  o[0] = c->time; o[1] = c->dt; o[2] = c->beta; o[3] = c->al;
  o[4] = c->ga; o[5] = c->ro; o[6] = c->vinner_t; o[7] = c->dvinner_t;
  return 0;
}

int tc_begin_step(tc_ctx *c) {
  CTXCHK(c);
  return begin_step(c);
}

int tc_set_substep(tc_ctx *c, int ns) {
  CTXCHK(c);
  if (ns < 1 || ns > c->p.nsst) return fail("substep out of range");
  set_substep(c, ns);
  return 0;
}

int tc_set_fused(tc_ctx *c, int fused) {
  CTXCHK(c);
  c->fused = fused != 0;
  return 0;
}

#define ROUTINE(name, body) \
  int name(tc_ctx *c) {     \
    CTXCHK(c);              \
    c->cfl_valid = false;   \
    body;                   \
    return launch_status(); \
  }

ROUTINE(tc_explicit_terms_qt, tc::launch_explicit_qt(c))
ROUTINE(tc_explicit_terms_qr, tc::launch_explicit_qr(c))
ROUTINE(tc_explicit_terms_qz, tc::launch_explicit_qz(c))
ROUTINE(tc_implicit_and_update_qt, tc::launch_implicit_qt(c))
ROUTINE(tc_implicit_and_update_qr, tc::launch_implicit_qr(c))
ROUTINE(tc_implicit_and_update_qz, tc::launch_implicit_qz(c))
ROUTINE(tc_calc_local_divergence, tc::launch_divergence(c, 1.0 / (c->al * c->dt)))
ROUTINE(tc_solve_pressure_correction, tc::launch_poisson(c))
ROUTINE(tc_correct_velocity, tc::launch_correct_velocity(c, c->al))
ROUTINE(tc_correct_pressure, tc::launch_correct_pressure(c))

int tc_time_marcher(tc_ctx *c) {
  CTXCHK(c);
  c->cfl_valid = false;
  int rc = begin_step(c);
  if (rc) return rc;
  for (int ns = 1; ns <= c->p.nsst; ns++) {
    set_substep(c, ns);
    rc = run_substep(c);
    if (rc) return rc;
  }
  return 0;
}

int tc_project(tc_ctx *c) {
  CTXCHK(c);
  c->cfl_valid = false;
  const double al = c->al, dt = c->dt;
  c->al = 1.0;
  c->dt = 1.0;
  tc::launch_divergence(c, 1.0);
  tc::launch_poisson(c);
  tc::launch_correct_velocity(c, 1.0);
  c->al = al;
  c->dt = dt;
  int rc = launch_status();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int tc_calc_max_cfl(tc_ctx *c, double *cflm) {
  CTXCHK(c);
  tc::launch_max_cfl(c);
  int rc = launch_status();
  if (rc) return rc;
  double v;
  rc = read_reduction(c, 1, &v);
  if (rc) return rc;
  *cflm = std::max(0.00000001, v);
  return 0;
}

int tc_check_divergence(tc_ctx *c, double *qmax) {
  CTXCHK(c);
  tc::launch_max_div(c);
  int rc = launch_status();
  if (rc) return rc;
  return read_reduction(c, 1, qmax);
}

int tc_check_divergence_noalloc(tc_ctx *c, double *qmax_dev) {
  CTXCHK(c);
  if (!qmax_dev) return fail("tc_check_divergence_noalloc: null output");
  return tc::check_divergence_noalloc(c, qmax_dev);
}

int tc_calc_torque(tc_ctx *c, double *inner, double *outer) {
  CTXCHK(c);
  tc::launch_torque(c);
  int rc = launch_status();
  if (rc) return rc;
  double s[2];
  rc = read_reduction(c, 2, s);
  if (rc) return rc;
  torque_from_sums(c, s[0], s[1], inner, outer);
  return 0;
}

int tc_calc_velocity_stats(tc_ctx *c, double *vmax3, double *rms3) {
  CTXCHK(c);
  tc::launch_velocity_stats(c);
  int rc = launch_status();
  if (rc) return rc;
  double v[6];
  rc = read_reduction(c, 6, v);
  if (rc) return rc;
  const double volt = c->g.pi * (c->p.rext * c->p.rext - c->p.rint * c->p.rint);
  const double den = (double)c->nthm * (double)c->nzm * volt;
  for (int q = 0; q < 3; q++) {
    vmax3[q] = v[q];
    rms3[q] = v[3 + q] / den;
  }
  return 0;
}

static int run_stats(tc_ctx *c, bool stats, bool diss);

int tc_step(tc_ctx *c, int ntime, double walltime, double *o) {
  CTXCHK(c);
  // main.F90:156-223
  int errorcode = 0;
  double instcfl;
  int rc = 0;
  if (c->cfl_valid) {
    instcfl = c->cfl_cached;  // same fields as the previous logging pass
  } else {
    rc = tc_calc_max_cfl(c, &instcfl);
    if (rc) return rc;
  }
  const tc_params &p = c->p;
  if (p.idtv != 0) {
    if (ntime > 1) {
      c->dt = (instcfl < 1.0e-8) ? p.dtmax : p.limitcfl / instcfl;
      if (c->dt > p.dtmax) c->dt = p.dtmax;
    } else {
      c->dt = p.dtmin;
    }
    if (c->dt < p.dtmin) errorcode = 166;
  } else {
    c->dt = p.dtmax;
    instcfl = instcfl * c->dt;
    if (instcfl > p.limitcfl) errorcode = 165;
  }
  double out[9] = {instcfl, c->dt, 0, 0, 0, 0, 0, 0, 0};
  const double tnew = c->time + c->dt;
  const bool logs = (ntime == 1 || std::fmod(tnew, p.tout) < c->dt);
  // where the last correction pass is the fused one it also evaluates the
  // diagnostics: a logging iteration uses them all, any other keeps the
  // CFL for the next iteration's dt rule (CalcMaxCFL of main.F90:156 on
  // the same fields), which saves that iteration its own CFL pass
  c->diag_in_correction = c->fused && !c->batch_owned && tc::correct_diag_applies(c);
  rc = tc_time_marcher(c);
  const bool diag_done = c->diag_in_correction;
  c->diag_in_correction = false;
  if (rc) return rc;
  c->time = tnew;
  out[2] = c->time;
  if (logs) {
    // CalcVelocityStats, CalcMaxCFL, CheckDivergence, CalcTorque in one pass
    if (!diag_done) tc::launch_diagnostics(c);
    rc = launch_status();
    if (rc) return rc;
    double v[10];
    rc = read_reduction(c, 10, v);
    if (rc) return rc;
    if (v[2] > p.limitvel && v[3] > p.limitvel) errorcode = 266;
    out[0] = std::max(0.00000001, v[0]);
    c->cfl_cached = out[0];
    c->cfl_valid = true;
    if (p.idtv == 0) out[0] = out[0] * c->dt;  // main.F90:201
    out[3] = v[1];
    torque_from_sums(c, v[8], v[9], &out[4], &out[5]);
    out[6] = v[2];
    out[7] = v[3];
    out[8] = v[4];
    if (std::fabs(out[3]) > p.resid) errorcode = 169;
    if (c->time > p.tsta) {  // main.F90:188-190
      rc = run_stats(c, true, true);
      if (rc) return rc;
    }
  } else if (diag_done) {
    rc = launch_status();
    if (rc) return rc;
    double v[10];
    rc = read_reduction(c, 10, v);
    if (rc) return rc;
    c->cfl_cached = std::max(0.00000001, v[0]);
    c->cfl_valid = true;
  }
  if (c->time > p.tmax) errorcode = 333;
  if (walltime > p.walltimemax) errorcode = 334;
  if (ntime == p.ntst) errorcode = 555;
  if (o)
    for (int q = 0; q < 9; q++) o[q] = out[q];
  return errorcode;
}

// One statistics pass; with diss, nute = volume sum / (nthm nzm volt
// lamdissip) as CalcDissipation.F90:142-148 forms it.
static int run_stats(tc_ctx *c, bool stats, bool diss) {
  tc::launch_stats(c, stats, diss);
  int rc = launch_status();
  if (rc) return rc;
  if (stats) c->nstatsamples = c->nstatsamples + 1;
  if (diss) {
    double sum;
    rc = read_reduction(c, 1, &sum);
    if (rc) return rc;
    const tc_params &p = c->p;
    const double pi = c->g.pi;
    const double volt = pi * (p.rext * p.rext - p.rint * p.rint);
    const double w = (p.virotint / p.rint - p.virotext / p.rext) / (p.rext - p.rint);
    const double lamdissip =
        4.0 * (p.rext * p.rext) * (p.rint * p.rint) / ((p.rext + p.rint) * (p.rext + p.rint)) * (w * w);
    c->last_nute = sum / ((double)c->nthm * (double)c->nzm * volt * lamdissip);
  }
  return 0;
}

int tc_calc_stats(tc_ctx *c) {
  CTXCHK(c);
  return run_stats(c, true, false);
}

int tc_calc_dissipation(tc_ctx *c, double *nute) {
  CTXCHK(c);
  const int rc = run_stats(c, false, true);
  if (rc) return rc;
  if (nute) *nute = c->last_nute;
  return 0;
}

int tc_get_stats(tc_ctx *c, double *out, int *nsamples) {
  CTXCHK(c);
  if (out) {
    HIPCHK(hipMemcpyAsync(out, c->stats, sizeof(double) * TC_NSTATS * (size_t)c->nrm * c->nzm,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (nsamples) *nsamples = c->nstatsamples;
  return 0;
}

int tc_set_stats(tc_ctx *c, const double *in, int nsamples) {
  CTXCHK(c);
  if (!in) return fail("null stats array");
  if (nsamples < 0) return fail("negative sample count");
  HIPCHK(hipMemcpyAsync(c->stats, in, sizeof(double) * TC_NSTATS * (size_t)c->nrm * c->nzm,
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nstatsamples = nsamples;
  return 0;
}

// Slab dumps: extraction on the device, three arrays of n doubles out.
static int run_dump(tc_ctx *c, int kind, int k, double *vt, double *vr, double *vz, size_t n) {
  if (!vt || !vr || !vz) return fail("null output array");
  if (!c->dump) {
    const size_t m = std::max((size_t)c->P, (size_t)c->nrm * c->nzm);
    if (dev_alloc(c, (void **)&c->dump, sizeof(double) * 3 * m)) return fail("out of device memory");
  }
  tc::launch_dump(c, kind, k);
  int rc = launch_status();
  if (rc) return rc;
  double *outs[3] = {vt, vr, vz};
  for (int q = 0; q < 3; q++)
    HIPCHK(hipMemcpyAsync(outs[q], c->dump + q * n, sizeof(double) * n, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int tc_dump_rcut(tc_ctx *c, int kslab, double *vt, double *vr, double *vz) {
  CTXCHK(c);
  if (kslab < 1 || kslab > c->nrm) return fail("kslab out of 1..nrm");
  return run_dump(c, 0, kslab, vt, vr, vz, (size_t)c->P);
}

int tc_dump_thcut(tc_ctx *c, double *vt, double *vr, double *vz) {
  CTXCHK(c);
  return run_dump(c, 1, 0, vt, vr, vz, (size_t)c->nrm * c->nzm);
}

int tc_dump_streamwise_avg(tc_ctx *c, double *vt, double *vr, double *vz) {
  CTXCHK(c);
  return run_dump(c, 2, 0, vt, vr, vz, (size_t)c->nrm * c->nzm);
}

int tc_last_dissipation(tc_ctx *c, double *nute) {
  CTXCHK(c);
  if (nute) *nute = c->last_nute;
  return 0;
}

// main.F90:152-261.  The wall clock is read after each iteration's work,
// as main.F90:208,220 reads ti(2) after TimeMarcher and the logging pass:
// WALLTIMEMAX then replaces that iteration's code with 334, except 555,
// which main.F90 assigns last.
static int run_steps(tc_ctx *c, int ntime0, int nsteps, double walltime_start, int *ntime_last,
                     double *o, double *log10, int log_cap, int *nlog) {
  CTXCHK(c);
  if (nsteps < 1) return fail("tc_run_steps needs nsteps >= 1");
  if (log10 && log_cap < 0) return fail("negative log capacity");
  const auto t0 = std::chrono::steady_clock::now();
  const tc_params &p = c->p;
  int code = 0, n = ntime0, nl = 0;
  double o9[9] = {0};
  for (int q = 0; q < nsteps; q++) {
    n = ntime0 + q;
    code = tc_step(c, n, -HUGE_VAL, o9);
    if (code < 0) break;
    const double wall =
        walltime_start + std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (wall > p.walltimemax && code != 555) code = 334;
    // the logging iterations of main.F90:180 (tc_step logs on the same rule)
    if (log10 && (n == 1 || std::fmod(c->time, p.tout) < c->dt)) {
      if (nl < log_cap) {
        log10[10 * (size_t)nl] = (double)n;
        for (int v = 0; v < 9; v++) log10[10 * (size_t)nl + 1 + v] = o9[v];
      }
      nl++;
    }
    if (code != 0) break;
  }
  if (ntime_last) *ntime_last = n;
  if (o)
    for (int v = 0; v < 9; v++) o[v] = o9[v];
  if (nlog) *nlog = nl;
  return code;
}

int tc_run_steps(tc_ctx *c, int ntime0, int nsteps, double walltime_start, int *ntime_last,
                 double *o) {
  return run_steps(c, ntime0, nsteps, walltime_start, ntime_last, o, nullptr, 0, nullptr);
}

int tc_run_steps_logged(tc_ctx *c, int ntime0, int nsteps, double walltime_start, int *ntime_last,
                        double *o, double *log10, int log_cap, int *nlog) {
  return run_steps(c, ntime0, nsteps, walltime_start, ntime_last, o, log10, log_cap, nlog);
}

int tc_init_perturbed_couette(const tc_params *p, unsigned long long seed, double amp, double *qt,
                              double *qr, double *qz) {
  if (!p || !qt || !qr || !qz) return fail("null argument");
  tc::Grid g;
  std::string e = tc::build_grid(*p, g);
  if (!e.empty()) return fail(e);
  const int nr = g.nr, nrm = g.nrm;
  const size_t n = (size_t)nr * p->nzm * p->nthm;
  unsigned long long state = seed;
  auto next = [&state]() {
    unsigned long long z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  // CreateInitialConditions.F90:21-23 laminar profile coefficients
  const double eta = p->rint / p->rext;
  const double alam = (p->virotext - eta * eta * p->virotint) / (1 - eta * eta);
  const double blam = (p->virotint - p->virotext) * p->rint * p->rint / (1 - eta * eta);
  for (size_t q = 0; q < n; q++) {
    const int k = (int)(q % nr) + 1;
    qt[q] = 0.0;
    if (k <= nrm) qt[q] = alam * g.rm[k] + blam / g.rm[k] + amp * (2.0 * next() - 1.0);
  }
  for (size_t q = 0; q < n; q++) {
    const int k = (int)(q % nr) + 1;
    qr[q] = 0.0;
    if (k >= 2 && k <= nrm) qr[q] = amp * (2.0 * next() - 1.0);
  }
  for (size_t q = 0; q < n; q++) {
    const int k = (int)(q % nr) + 1;
    qz[q] = 0.0;
    if (k <= nrm) qz[q] = amp * (2.0 * next() - 1.0);
  }
  return 0;
}

int tc_synchronize(tc_ctx *c) {
  CTXCHK(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// Events are taken from a pool; creating them while the stream is idle at a
// step boundary would sit on the critical path, so the pool is filled here.
static void fill_event_pool(tc_ctx *c, size_t n) {
  while (c->event_pool.size() < n) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) break;
    c->event_pool.push_back(e);
  }
}

int tc_kernel_timing(tc_ctx *c, int enable) {
  CTXCHK(c);
  c->timing = enable != 0;
  c->timing_only.clear();
  if (c->timing) fill_event_pool(c, 1024);
  return 0;
}

int tc_kernel_timing_only(tc_ctx *c, const char *kernel) {
  CTXCHK(c);
  c->timing = kernel != nullptr;
  c->timing_only = kernel ? kernel : "";
  if (c->timing) fill_event_pool(c, 1024);
  return 0;
}

int tc_kernel_timing_read(tc_ctx *c, const char *kernel, double *total_ms, long long *count,
                          int reset) {
  CTXCHK(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &t : c->timers) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
    auto &acc = c->timing_acc[t.name];
    acc.first += ms;
    acc.second += 1;
    c->event_pool.push_back(t.a);
    c->event_pool.push_back(t.b);
  }
  c->timers.clear();
  auto it = c->timing_acc.find(kernel ? kernel : "");
  if (total_ms) *total_ms = (it == c->timing_acc.end()) ? 0.0 : it->second.first;
  if (count) *count = (it == c->timing_acc.end()) ? 0 : it->second.second;
  if (reset) c->timing_acc.clear();
  return 0;
}

int tc_copy_field(tc_ctx *c, int dst, int src) {
  CTXCHK(c);
  if (check_field(c, dst) || check_field(c, src)) return -1;
  if (dst == src) return fail("tc_copy_field needs two different fields");
  c->cfl_valid = false;
  tc::launch_copy(c, dst, src);
  return launch_status();
}

int tc_stream_triad(tc_ctx *c, int a, int b, int cc, double s) {
  CTXCHK(c);
  if (check_field(c, a) || check_field(c, b) || check_field(c, cc)) return -1;
  if (a == b || a == cc || b == cc) return fail("tc_stream_triad needs three different fields");
  c->cfl_valid = false;
  tc::launch_triad(c, a, b, cc, s);
  return launch_status();
}

int tc_device_bytes(tc_ctx *c, unsigned long long *bytes) {
  CTXCHK(c);
  *bytes = c->device_bytes;
  return 0;
}

}  // extern "C"
