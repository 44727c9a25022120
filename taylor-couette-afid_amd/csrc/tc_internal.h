// Internal definitions shared by the MI355X kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/tc_abi.h"
#include "tc_grid.h"

namespace tc {

// Device view of the grid: pointers into one device buffer of 1-based
// vectors (v[k] == reference v(k)) plus the scalars the kernels need.
// The radial coefficients of the explicit terms and right-hand sides at one
// radial index k, gathered into one 192-byte record so a kernel marching
// along r fetches them with three scalar loads per plane from one pointer.
// Each entry is evaluated on the host with the expression (and operand
// order) the kernels used to evaluate on the device.
struct HelmPlane {
  double udxt_m;   // dxt*0.25*usrm(k)
  double udxtq_m;  // dxtq/(ren*(rm(k)*rm(k)))
  double udx2rm, usrm, usrc, usrc_p, udx1vis1;  // usrc_p = usrc(k+1)
  double udxt_c;   // dxt*0.25*usrc(k)
  double udxtq_c;  // dxtq/(ren*(rc(k)*rc(k)))
  double usrc_m, udx2c, udx1vis2, rc;           // usrc_m = usrc(k-1)
  double ap1j, ac1j, am1j, ap2je, ac2je, am2je, ap3j, ac3j, am3j;
  double udxzq;    // dxzq/ren
  double pad;
};
static_assert(sizeof(HelmPlane) == 192, "HelmPlane is three 64-byte loads");

struct GridDev {
  const HelmPlane *hp;  // [k], k = 1..nrm
  const double *rc, *rm, *usrc, *usrm, *udx2rm, *udx2m, *udx2c, *rudx2c, *udx1vis1,
      *udx1vis2, *udx1qm, *pvol;
  const double *ap1j, *am1j, *ac1j, *ap2je, *am2je, *ac2je, *ap3j, *am3j, *ac3j;
  const double *apph, *amph, *acph, *ak1, *ak3;
  const double *udx2mm, *udx2cc;  // statistics and dissipation only
  double dxt, dxz, dxtq, dxzq, ren, romeg;
  double ap1j_fs2, ac1j_fs1, ac1j_fs2, am1j_fs1, ap3j_fs2, ac3j_fs1, ac3j_fs2, am3j_fs1;
  int nthm, nzm, nrm, nr, nth2, sstr;  // sstr: spectral row stride >= nth2
  long long P;  // points per r-plane = nzm*nthm
};

// Radix plan of one FFT length.
struct FFTPlan {
  int n = 0;
  int nst = 0;
  int radix[24] = {0};
  int lines = 1;  // FFT lines per 256-thread block
  int me = 8;     // points per thread per pass (n <= me * nt/lines)
  int nt = 256;   // threads per block
  int padsh = 3;  // one LDS padding slot per 2^padsh points (31: none)
};

// Helmholtz LU tables: per (substep slot, component, mask variant) a block of
// NT rows of LU_ROW doubles: dl, du, du2, rd=1/d, row scale, pivot flag, d, 0
// (d, du, du2 are the rows of the upper factor U).
// One row is one 64-byte scalar load.
enum { LU_QT = 0, LU_QR = 1, LU_QZ = 2 };
constexpr int LU_ROW = 8;

struct Timer {
  std::string name;
  hipEvent_t a, b;
};

// Scalars of one substep's implicit solves
struct ImpArgs {
  double al, ga, ro, dt, alre;
  double vwall_in;  // virotint + vinner_t
  double virotext;
  double bc1;       // 0.5*al*dt*dvinner_t*am1j(1)/ren (SolveImpEqnUpdate_QT.F90:67)
  double udxz;      // dxz*al (ImplicitAndUpdateQZ.F90:25)
};

// Everything a grouped launch needs about one instance for one substep; one
// table of these per (shape group, substep) lives in device memory.
struct StepDesc {
  GridDev g;
  ImpArgs a;                    // vwall_in and bc1 are overridden from BatchDyn
  double *f[TC_NFIELDS];
  double2 *spec;
  const unsigned char *icyl, *ocyl;
  const double *lut, *lur, *luz;  // LU tables of this substep
  int lu_nt;
  double dt, al, beta, usdt, inv_norm;
  FFTPlan pt, pz;
  const double2 *twt, *twz;
  double *red, *red_out;        // reduction partials and results
  double limitcfl, resid, limitvel;
  double virotint, virotext, del, deln, pi;
  // deferred pressure correction (corr != 0): the Helmholtz kernel applies
  // the previous substep's correction (al corr_al) reading pr_in and writing
  // pr_out; otherwise pr_in is the pressure field
  int corr;
  double corr_al;
  const double *pr_in;
  double *pr_out;
  // this instance's step scalars and device results (its exit state)
  const struct BatchDyn *dyn;
  const struct BatchRes *res;
};

// Per-instance scalars that change every step (written by the host before
// each step): the inner-wall terms of the three substeps, whether this step
// logs, and the step index.
struct BatchDyn {
  double vwall[3], bc1[3];
  double tnew;  // time at the end of this step
  int log, ntime;
};

// Per-instance results kept on the device between host synchronisations.
struct BatchRes {
  double cfl;          // CFL at the start of the last step
  double diag[10];     // last logging pass: cfl, max|div|, vmax(3), rms sums(3), torque sums(2)
  int fail165;         // 1 + ntime of the first step with cfl*dt > CFLMAX (0: none)
  int log_ntime;       // ntime of the last logging pass (-1: none)
  int faillog;         // 1 + ntime of the first logging step with code 169 or 266 (0: none)
  int faillog_code;    // that code
  double fail_time;    // time at the end of the first failing step
};

// An instance whose run has ended on the device (code 165 at the start of a
// step, or 169/266 in its logging pass) completes that step, as the main
// loop does, and takes no further step: every grouped kernel leaves at once
// for it from the next step on.
#ifdef __HIPCC__
// A pointer read from a descriptor table in device memory is generic to the
// compiler, which then addresses it with flat instructions and 64-bit VGPR
// addresses (about 50 more VGPRs and one wave less per SIMD in the grouped
// Helmholtz kernel than in its plain twin, whose pointers are kernel
// arguments).  Every pointer in a StepDesc is device memory; passing it
// through an integer into the global address space gives global loads with
// scalar bases, as kernel arguments get.
template <class T>
__device__ __forceinline__ T *glob(T *p) {
  typedef __attribute__((address_space(1))) T GT;
  return (T *)(GT *)(unsigned long long)p;
}
__device__ __forceinline__ GridDev glob(const GridDev &s) {
  GridDev g = s;
  g.hp = glob(g.hp);
  g.rc = glob(g.rc), g.rm = glob(g.rm), g.usrc = glob(g.usrc), g.usrm = glob(g.usrm);
  g.udx2rm = glob(g.udx2rm), g.udx2m = glob(g.udx2m), g.udx2c = glob(g.udx2c);
  g.rudx2c = glob(g.rudx2c), g.udx1vis1 = glob(g.udx1vis1), g.udx1vis2 = glob(g.udx1vis2);
  g.udx1qm = glob(g.udx1qm), g.pvol = glob(g.pvol);
  g.ap1j = glob(g.ap1j), g.am1j = glob(g.am1j), g.ac1j = glob(g.ac1j);
  g.ap2je = glob(g.ap2je), g.am2je = glob(g.am2je), g.ac2je = glob(g.ac2je);
  g.ap3j = glob(g.ap3j), g.am3j = glob(g.am3j), g.ac3j = glob(g.ac3j);
  g.apph = glob(g.apph), g.amph = glob(g.amph), g.acph = glob(g.acph);
  g.ak1 = glob(g.ak1), g.ak3 = glob(g.ak3), g.udx2mm = glob(g.udx2mm), g.udx2cc = glob(g.udx2cc);
  return g;
}
__device__ __forceinline__ bool inst_stopped(const StepDesc &d) {
  if (!d.res) return false;
  const BatchRes *r = glob(d.res);
  const int a = r->fail165, b = r->faillog;
  const int f = (a && b) ? (a < b ? a : b) : (a ? a : b);
  return f != 0 && glob(d.dyn)->ntime >= f;
}
#endif

}  // namespace tc

struct tc_ctx {
  tc_params p;
  tc::Grid g;
  int nthm = 0, nzm = 0, nrm = 0, nr = 0, nth2 = 0, sstr = 0;
  long long P = 0, N = 0;  // plane size, field size (nr planes)
  hipStream_t stream = nullptr;
  int device = 0;
  double *f[TC_NFIELDS] = {nullptr};
  double2 *spec = nullptr;   // nrm x nzm x sstr complex, [r][kz][ktheta]
  double *gvbuf = nullptr;   // device grid vectors
  tc::GridDev gd;
  unsigned char *mask = nullptr;  // device [2][nzm*nthm] (icyl, ocyl), [j][i]
  std::vector<int> icyl, ocyl;    // host copies, (nzm,nthm) column-major
  bool masks_uniform_one = true;
  // Helmholtz LU tables
  double *lu = nullptr;        // device
  double *lu_host = nullptr;   // pinned staging
  hipEvent_t lu_copied = nullptr;
  int lu_nt = 0;               // doubles per table = LU_ROW*(nrm+1)
  double lu_key[3] = {-1, -1, -1};
  bool lu_valid = false;
  bool lu_all_variants = false;  // all four mask variants built (else only variant 3)
  // FFT
  tc::FFTPlan plan_t, plan_z;
  double2 *tw_t = nullptr, *tw_z = nullptr;
  // reductions
  double *red = nullptr;        // device partials
  size_t red_cap = 0;           // their capacity in doubles
  double *red_out = nullptr;    // device results (16)
  double *red_host = nullptr;   // pinned copy of red_out
  int red_blocks = 0;
  // scalars of module param
  double time = 0, dt = 0, beta = 0, al = 0, ga = 0, ro = 0, vinner_t = 0, dvinner_t = 0;
  int ns = 1;
  bool fused = true;
  // fused RK3 step: the pressure correction of substeps 1 and 2 is applied
  // by the next substep's Helmholtz kernel instead of its own pass; pr_alt
  // receives the corrected pressure of substep 1 (TC_DEFER_CORRECTION at
  // tc_create: 1 from 8 Mi grid points up, 2 always, 0 never)
  bool defer_corr = true;
  double corr_al = 0.0;
  double *pr_alt = nullptr;
  int divfft_stage = 1;  // fused divergence FFT: stage the divergence rows in LDS first
  int poisson_pq = 13;  // rows per thread of the radial Poisson solve (13 or 26)
  int helm_lpt = 1;  // radial lines per thread in the fused Helmholtz kernel (1 or 2)
  tc_tuning tuning{};  // effective tuning (helm_lpt 0 = chosen by occupancy)
  // statistics accumulators (TC_NSTATS x nrm x nzm) and per-line partials
  double *stats = nullptr, *stat_part = nullptr;
  int nstatsamples = 0;
  double last_nute = 0.0;  // NaN until the first CalcDissipation
  double *dump = nullptr;  // slab extraction, 3 x max(P, nrm*nzm), made at first use
  // CFL of the current velocity, when known (set by the logging pass of tc_step)
  bool cfl_valid = false;
  // tc_step of a logging iteration: the last substep's correction pass also
  // evaluates the logging diagnostics (launch_correct_diag)
  bool diag_in_correction = false;
  bool batch_owned = false;     // advanced by a tc_batch (its tables hold the field pointers)
  double cfl_cached = 0.0;
  // timing
  bool timing = false;
  bool timer_open = false;   // the last timer_begin recorded a start event
  std::string timing_only;   // when set, only this kernel name is timed
  std::vector<tc::Timer> timers;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, long long>> timing_acc;
  unsigned long long device_bytes = 0;
};

namespace tc {

// kernel launchers (tc_kernels.hip)
void launch_explicit_qt(tc_ctx *c);
void launch_explicit_qr(tc_ctx *c);
void launch_explicit_qz(tc_ctx *c);
void launch_helmholtz_fused(tc_ctx *c);
void launch_helmholtz_fused_corr(tc_ctx *c, const double *pr_in, double *pr_out, double al_prev);
int choose_helm_lpt(const tc_ctx *c);
void launch_implicit_qt(tc_ctx *c);
void launch_implicit_qr(tc_ctx *c);
void launch_implicit_qz(tc_ctx *c);
void launch_divergence(tc_ctx *c, double usdt);
void launch_correct_velocity(tc_ctx *c, double aldt);
void launch_correct_pressure(tc_ctx *c);
void launch_correct_both(tc_ctx *c);
void launch_zero_wall_qr(tc_ctx *c);
// reductions: results land in c->red_out and are copied to c->red_host
void launch_max_cfl(tc_ctx *c);
void launch_max_div(tc_ctx *c);
void launch_torque(tc_ctx *c);
void launch_velocity_stats(tc_ctx *c);
// one pass: cfl, max|div|, vmax(3), rms sums(3), torque sums(2)
void launch_diagnostics(tc_ctx *c);
bool correct_diag_applies(const tc_ctx *c);
int check_divergence_noalloc(tc_ctx *c, double *qmax_dev);
void launch_correct_diag(tc_ctx *c);
void launch_stats(tc_ctx *c, bool stats, bool diss);
void launch_dump(tc_ctx *c, int kind, int k);
int regrid_field(tc_ctx *c, int field, const double *old, int ntho, int nzo, int nro, int istro,
                 double stro);
void launch_copy(tc_ctx *c, int dst, int src);
void launch_triad(tc_ctx *c, int a, int b, int cc, double s);

// grouped launches over n instances of one shape (c: any instance of the
// group, for the shape and the tuning); tab: device table of n StepDesc of
// the substep, dyn/res: device arrays of n
bool defers_correction(const tc_ctx *c);
int set_deferral(tc_ctx *c, int mode);
// the dt and step recorded in a restart file's header (nothing else is read)
int restart_file_info(const char *path, double *dt, int *ntime);
void launch_helmholtz_fused_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, int ns, int n,
                              hipStream_t s, bool corr);
void launch_correct_both_g(tc_ctx *c, const StepDesc *tab, int n, hipStream_t s);
void launch_cfl_check_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, BatchRes *res, int n,
                        hipStream_t s);
void launch_diagnostics_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, BatchRes *res, int n,
                          hipStream_t s);
void launch_poisson_g(tc_ctx *c, const StepDesc *tab, int n, hipStream_t s);

// Poisson solve (tc_poisson.hip)
std::string make_fft_plan(int n, FFTPlan &pl, int lds_bytes, int nt);
void fill_twiddles(int n, std::vector<double> &tw);
// with_divergence: compute dph = div(q)*usdt inside the forward theta FFT
// instead of reading it (the fused time marcher)
void launch_poisson(tc_ctx *c, bool with_divergence = false, double usdt = 0.0);

// context helpers shared with the batch layer (tc_abi.cpp)
int ctx_begin_step(tc_ctx *c);
void ctx_set_substep(tc_ctx *c, int ns);
void ctx_torque(const tc_ctx *c, double slow, double sup, double *inner, double *outer);
int set_error(const std::string &msg);
int check_launch();

// timing helpers (tc_abi.cpp)
void timer_begin(tc_ctx *c, const char *name);
void timer_end(tc_ctx *c);

}  // namespace tc
