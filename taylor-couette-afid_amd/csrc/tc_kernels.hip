// MI355X (gfx950) kernels for the stencil, Helmholtz and reduction parts of
// the Taylor-Couette RK3 step.  FP64 throughout, wave64.
//
// Device layout: every 3-D field is [r][z][theta] with theta contiguous,
// r-plane k (1-based, k = 1..nr) at offset (k-1)*P, P = nzm*nthm.  The
// reference's periodic halos (InitVariables.F90:26-35, halo_common.f90:67-108)
// become index wrap-around, so no halo update is ever launched.
#include <map>
#include <mutex>

#include "tc_internal.h"

namespace tc {

#define DEVI __device__ __forceinline__

DEVI long long IX(const GridDev &g, int k, int j, int i) {
  return (long long)(k - 1) * g.P + (long long)j * g.nthm + i;
}
DEVI int wrap_p(int i, int n) { return (i + 1 == n) ? 0 : i + 1; }
DEVI int wrap_m(int i, int n) { return (i == 0) ? n - 1 : i - 1; }

// Non-temporal access for arrays streamed once per kernel (no other block
// re-reads them soon): a streaming copy runs 5.8% faster with both.
DEVI double ld_nt(const double *p) { return __builtin_nontemporal_load(p); }
DEVI void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
DEVI double2 ld2_nt(const double *p) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
  return make_double2(__builtin_nontemporal_load(&q->x), __builtin_nontemporal_load(&q->y));
}
DEVI void st2_nt(double *p, double u, double v) {
  double2 *q = reinterpret_cast<double2 *>(p);
  __builtin_nontemporal_store(u, &q->x);
  __builtin_nontemporal_store(v, &q->y);
}

// ---------------------------------------------------------------------------
// Explicit terms.  Written once, generic over an accessor A with
// A(dk, dj, di) = field value at (k+dk, j+dj, i+di) of the point being
// computed, so the routine-level kernels (global loads) and the fused
// Helmholtz kernel (LDS tiles) evaluate identical arithmetic.
// ---------------------------------------------------------------------------

// Global-memory accessor with periodic wrap in z and theta.
struct GAcc {
  const double *__restrict__ a;
  long long base;  // IX(k, j, i)
  long long P;
  int i, j, nthm, nzm;
  DEVI double operator()(int dk, int dj, int di) const {
    int jj = j + dj, ii = i + di;
    jj = (jj < 0) ? jj + nzm : (jj >= nzm ? jj - nzm : jj);
    ii = (ii < 0) ? ii + nthm : (ii >= nthm ? ii - nthm : ii);
    return a[base + dk * P + (long long)(jj - j) * nthm + (ii - i)];
  }
};
DEVI GAcc gacc(const GridDev &g, const double *a, int k, int j, int i) {
  GAcc r;
  r.a = a;
  r.base = IX(g, k, j, i);
  r.P = g.P;
  r.i = i;
  r.j = j;
  r.nthm = g.nthm;
  r.nzm = g.nzm;
  return r;
}

// ExplicitTermsQT.F90:44-125 at (kc, j, i)
template <class A>
DEVI double explicit_qt_g(const GridDev &g, const HelmPlane &c, const A &T, const A &R,
                          const A &Z, int kc) {
  const int dkm = kc > 1 ? -1 : 0, dkp = kc < g.nrm ? 1 : 0;  // kmv, kpv clamps
  const double udxt = c.udxt_m;
  const double udxtq = c.udxtq_m;
  const double udxzq = c.udxzq;
  const double udxz = g.dxz * 0.25;
  const double t_c = T(0, 0, 0), t_ip = T(0, 0, 1), t_im = T(0, 0, -1);
  const double t_jp = T(0, 1, 0), t_jm = T(0, -1, 0);
  const double t_kp = T(dkp, 0, 0), t_km = T(dkm, 0, 0);
  const double r_c = R(0, 0, 0), r_im = R(0, 0, -1);
  const double r_kp = R(1, 0, 0), r_kpim = R(1, 0, -1);
  const double z_c = Z(0, 0, 0), z_im = Z(0, 0, -1);
  const double z_jp = Z(0, 1, 0), z_jpim = Z(0, 1, -1);
  const double usrc_kp = c.usrc_p, usrc_k = c.usrc;

  const double h11 = ((t_ip + t_c) * (t_ip + t_c) - (t_c + t_im) * (t_c + t_im)) * udxt;
  const double h12d = ((r_kp + r_kpim) * (t_kp + t_c) - (r_c + r_im) * (t_c + t_km)) * 0.25 * c.udx2rm;
  const double h12n = t_c * c.usrm * ((r_kp + r_kpim) * usrc_kp + (r_c + r_im) * usrc_k) * 0.25;
  const double h12 = h12d + h12n;
  const double h13 = ((z_jp + z_jpim) * (t_jp + t_c) - (z_c + z_im) * (t_c + t_jm)) * udxz;
  const double d11qt = (t_ip - 2.0 * t_c + t_im) * udxtq;
  const double d22qt = (t_jp - 2.0 * t_c + t_jm) * udxzq;
  const double qre = r_kp * usrc_kp + r_c * usrc_k;
  const double qrw = r_kpim * usrc_kp + r_im * usrc_k;
  const double d11qre = (qre - qrw) * c.udx1vis1 * 0.5;
  const double coriol = -((r_c + r_im) * usrc_k + (r_kp + r_kpim) * usrc_kp) * 0.25 * g.romeg;
  return -(h11 + h12 + h13) + d11qre + d22qt + d11qt + coriol;
}

// ExplicitTermsQR.F90:46-146 at (kc, j, i), kc = 2..nrm
template <class A>
DEVI double explicit_qr_g(const GridDev &g, const HelmPlane &c, const A &T, const A &R,
                          const A &Z, int kc) {
  const double udxt = c.udxt_c;
  const double udxtq = c.udxtq_c;
  const double udxzq = c.udxzq;
  const double udxz = g.dxz * 0.25;
  const double t_c = T(0, 0, 0), t_ip = T(0, 0, 1);
  const double t_m = T(-1, 0, 0), t_mip = T(-1, 0, 1);
  const double r_c = R(0, 0, 0), r_ip = R(0, 0, 1), r_im = R(0, 0, -1);
  const double r_jp = R(0, 1, 0), r_jm = R(0, -1, 0);
  const double r_kp = R(1, 0, 0), r_km = R(-1, 0, 0);
  const double z_c = Z(0, 0, 0), z_jp = Z(0, 1, 0);
  const double z_m = Z(-1, 0, 0), z_mjp = Z(-1, 1, 0);

  const double tavg = (t_ip + t_mip + t_c + t_m) * 0.25;
  const double h11n = tavg * tavg;
  const double h21 = ((t_ip + t_mip) * (r_ip + r_c) - (t_c + t_m) * (r_c + r_im)) * udxt;
  const double h22 = ((r_kp * c.usrc_p + r_c * c.usrc) * (r_kp + r_c) -
                      (r_c * c.usrc + r_km * c.usrc_m) * (r_c + r_km)) *
                     c.udx2c * 0.25;
  const double h23 = ((z_jp + z_mjp) * (r_jp + r_c) - (z_c + z_m) * (r_c + r_jm)) * udxz;
  const double qte = (t_ip + t_mip);
  const double qtw = (t_c + t_m);
  const double d11qre = -(qte - qtw) * c.udx1vis2;
  const double d11qr = (r_ip - 2.0 * r_c + r_im) * udxtq;
  const double d22qr = (r_jp - 2.0 * r_c + r_jm) * udxzq;
  const double coriol = (t_c + t_ip + t_m + t_mip) * 0.25 * c.rc * g.romeg;
  return -(h21 + h22 + h23) + h11n + d11qr + d22qr + d11qre + coriol;
}

// ExplicitTermsQZ.F90:40-102 at (kc, j, i)
template <class A>
DEVI double explicit_qz_g(const GridDev &g, const HelmPlane &c, const A &T, const A &R,
                          const A &Z, int kc) {
  const int dkm = kc > 1 ? -1 : 0, dkp = kc < g.nrm ? 1 : 0;
  const double udxt = c.udxt_m;
  const double udxtq = c.udxtq_m;
  const double udxzq = c.udxzq;
  const double udxz = g.dxz * 0.25;
  const double t_c = T(0, 0, 0), t_ip = T(0, 0, 1);
  const double t_jm = T(0, -1, 0), t_jmip = T(0, -1, 1);
  const double r_c = R(0, 0, 0), r_jm = R(0, -1, 0);
  const double r_kp = R(1, 0, 0), r_kpjm = R(1, -1, 0);
  const double z_c = Z(0, 0, 0), z_ip = Z(0, 0, 1), z_im = Z(0, 0, -1);
  const double z_jp = Z(0, 1, 0), z_jm = Z(0, -1, 0);
  const double z_kp = Z(dkp, 0, 0), z_km = Z(dkm, 0, 0);

  const double h31 = ((t_ip + t_jmip) * (z_ip + z_c) - (t_c + t_jm) * (z_c + z_im)) * udxt;
  const double h32 = (((r_kp + r_kpjm) * (z_kp + z_c)) - ((r_c + r_jm) * (z_c + z_km))) * c.udx2rm * 0.25;
  const double h33 = ((z_jp + z_c) * (z_jp + z_c) - (z_c + z_jm) * (z_c + z_jm)) * udxz;
  const double dqz1 = (z_im - 2.0 * z_c + z_ip) * udxtq;
  const double dqz2 = (z_jm - 2.0 * z_c + z_jp) * udxzq;
  return -(h31 + h32 + h33) + dqz1 + dqz2;
}

template <int WHICH>
__global__ __launch_bounds__(256) void k_explicit(GridDev g, const double *__restrict__ qt,
                                                  const double *__restrict__ qr,
                                                  const double *__restrict__ qz,
                                                  double *__restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  const int kc = blockIdx.z + (WHICH == 1 ? 2 : 1);
  const GAcc T = gacc(g, qt, kc, j, i), R = gacc(g, qr, kc, j, i), Z = gacc(g, qz, kc, j, i);
  const HelmPlane c = g.hp[kc];
  double v;
  if (WHICH == 0)
    v = explicit_qt_g(g, c, T, R, Z, kc);
  else if (WHICH == 1)
    v = explicit_qr_g(g, c, T, R, Z, kc);
  else
    v = explicit_qz_g(g, c, T, R, Z, kc);
  out[IX(g, kc, j, i)] = v;
}

static dim3 point_grid(const tc_ctx *c, int planes) {
  return dim3((unsigned)((c->nthm + 63) / 64), (unsigned)((c->nzm + 3) / 4), (unsigned)planes);
}

void launch_explicit_qt(tc_ctx *c) {
  timer_begin(c, "explicit_qt");
  hipLaunchKernelGGL(k_explicit<0>, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_DQ]);
  timer_end(c);
}
void launch_explicit_qr(tc_ctx *c) {
  timer_begin(c, "explicit_qr");
  hipLaunchKernelGGL(k_explicit<1>, point_grid(c, c->nrm - 1), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_DPH]);
  timer_end(c);
}
void launch_explicit_qz(tc_ctx *c) {
  timer_begin(c, "explicit_qz");
  hipLaunchKernelGGL(k_explicit<2>, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_QCAP]);
  timer_end(c);
}

// ---------------------------------------------------------------------------
// Helmholtz solves.  The LU factors (dgttrf with partial pivoting of the
// row-normalised matrices of SolveImpEqnUpdate_*) do not depend on the line
// (j, i) except through the wall masks, so they come from a per-substep
// table with one entry per mask combination.  One radial line per lane;
// lanes run along theta so every plane access is coalesced.  The forward
// substitution is streamed while the right-hand side is produced; its output
// y is written to a scratch field and read back by the backward substitution.
// ---------------------------------------------------------------------------

struct LUView {
  const double *b;  // rows of LU_ROW doubles: dl du du2 rd sc pv
  DEVI double dl(int r) const { return b[LU_ROW * r]; }
  DEVI double du(int r) const { return b[LU_ROW * r + 1]; }
  DEVI double du2(int r) const { return b[LU_ROW * r + 2]; }
  DEVI double rd(int r) const { return b[LU_ROW * r + 3]; }
  DEVI double sc(int r) const { return b[LU_ROW * r + 4]; }
  DEVI double pv(int r) const { return b[LU_ROW * r + 5]; }
  DEVI double ud(int r) const { return b[LU_ROW * r + 6]; }
  // row r of U q for q_r, q_(r+1), q_(r+2)
  DEVI double uq(int r, double q0, double q1, double q2) const {
    return ud(r) * q0 + du(r) * q1 + du2(r) * q2;
  }
};
DEVI LUView lu_view(const double *base, int nt) {
  LUView v;
  v.b = base;
  return v;
}

// streamed L-solve step for row r (0-based): f is the right-hand side of row r+1
DEVI void lsolve_step(const LUView &T, int r, double f, double &carry, double &out) {
  if (T.pv(r) == 0.0) {
    out = carry;
    carry = f - T.dl(r) * carry;
  } else {
    out = f;
    carry = carry - T.dl(r) * f;
  }
}

// Right-hand side row k of the q_theta solve (ImplicitAndUpdateQT.F90:45-115);
// qm, q0, qp = qt(k-1), qt(k), qt(k+1) on the line.
// NOSLIP: every wall mask is 1, so the blends with the free-slip
// coefficients reduce to the no-slip ones (x*1 + y*0 == x for the nonzero
// wall coefficients).
template <bool NOSLIP = false>
DEVI double rhs_qt(const GridDev &g, const HelmPlane &c, const ImpArgs &a, int k, double qm,
                   double q0, double qp, double pr_c, double pr_im, double dqv, double ruold,
                   double fic, double foc) {
  double d33;
  if (k == 1) {
    const double app = c.ap1j;
    const double acc = NOSLIP ? c.ac1j : (c.ac1j * fic + g.ac1j_fs1 * (1 - fic));
    const double amm = NOSLIP ? c.am1j : (c.am1j * fic + g.am1j_fs1 * (1 - fic));
    d33 = qp * app + q0 * acc + a.vwall_in * amm;
  } else if (k < g.nrm) {
    d33 = qp * c.ap1j + q0 * c.ac1j + qm * c.am1j;
  } else {
    const double app = NOSLIP ? c.ap1j : (c.ap1j * foc + g.ap1j_fs2 * (1 - foc));
    const double acc = NOSLIP ? c.ac1j : (c.ac1j * foc + g.ac1j_fs2 * (1 - foc));
    d33 = a.virotext * app + q0 * acc + qm * c.am1j;
  }
  const double dpx = (pr_c - pr_im) * g.dxt * a.al * c.usrm;
  return (a.ga * dqv + a.ro * ruold + a.alre * d33 - dpx) * a.dt;
}

// Right-hand side row k = 2..nrm of the q_r solve (ImplicitAndUpdateQR.F90:41-63).
DEVI double rhs_qr(const HelmPlane &c, const ImpArgs &a, double qm, double q0, double qp,
                   double pr_c, double pr_km, double dphv, double ruold) {
  const double d33 = qp * c.ap2je + q0 * c.ac2je + qm * c.am2je;
  const double udx2 = c.rc * a.al * c.udx2c;
  const double dpx = (pr_c - pr_km) * udx2;
  return (a.ga * dphv + a.ro * ruold + a.alre * d33 - dpx) * a.dt;
}

// Right-hand side row k of the q_z solve (ImplicitAndUpdateQZ.F90:42-113).
// Row nrm is the one the reference leaves behind: assembled a second time
// after the loop overwrote ru3(nrm) with qcap(nrm), so its history term is
// ro*qcap(nrm), with the loop's amm = am3j(nrm).
template <bool NOSLIP = false>
DEVI double rhs_qz(const GridDev &g, const HelmPlane &c, const ImpArgs &a, int k, double qm,
                   double q0, double qp, double pr_c, double pr_jm, double qcv, double ruold,
                   double fic, double foc) {
  double d33, hist = ruold;
  if (k == 1) {
    const double app = c.ap3j;
    const double acc = NOSLIP ? c.ac3j : (c.ac3j * fic + g.ac3j_fs1 * (1 - fic));
    const double amm = NOSLIP ? c.am3j : (c.am3j * fic + g.am3j_fs1 * (1 - fic));
    d33 = qp * app + q0 * acc + q0 * amm;
  } else if (k < g.nrm) {
    d33 = qp * c.ap3j + q0 * c.ac3j + qm * c.am3j;
  } else {
    const double app = NOSLIP ? c.ap3j : (c.ap3j * foc + g.ap3j_fs2 * (1 - foc));
    const double acc = NOSLIP ? c.ac3j : (c.ac3j * foc + g.ac3j_fs2 * (1 - foc));
    d33 = q0 * app + q0 * acc + qm * c.am3j;
    hist = qcv;
  }
  const double dpx = (pr_c - pr_jm) * a.udxz;
  return (a.ga * qcv + a.ro * hist + a.alre * d33 - dpx) * a.dt;
}

// ImplicitAndUpdateQT.F90:40-117 + SolveImpEqnUpdate_QT.F90:34-78
__global__ __launch_bounds__(256) void k_implicit_qt(GridDev g, ImpArgs a, double *__restrict__ qt,
                                                     const double *__restrict__ pr,
                                                     const double *__restrict__ dq,
                                                     double *__restrict__ ru1,
                                                     double *__restrict__ y,
                                                     const unsigned char *__restrict__ icyl,
                                                     const unsigned char *__restrict__ ocyl,
                                                     const double *__restrict__ lu, int nt) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.P) return;
  const int j = (int)(l / g.nthm), i = (int)(l - (long long)j * g.nthm);
  const long long lim = (long long)j * g.nthm + wrap_m(i, g.nthm);
  const int n = g.nrm;
  const int ic = icyl[l], oc = ocyl[l];
  const double fic = (double)ic, foc = (double)oc;
  const LUView T = lu_view(lu + (size_t)(ic + 2 * oc) * nt, nt);
  const long long P = g.P;
  double qm = 0.0, q0 = qt[l], qp = qt[l + P];
  double carry = 0.0, yprev;
  for (int k = 1; k <= n; k++) {
    const long long o = (long long)(k - 1) * P + l;
    if (k > 1) {
      qm = q0;
      q0 = qp;
      qp = (k < n) ? qt[o + P] : 0.0;
    }
    const double dqv = dq[o];
    const double r = rhs_qt(g, g.hp[k], a, k, qm, q0, qp, pr[o], pr[o - l + lim], dqv, ru1[o], fic, foc);
    ru1[o] = dqv;
    if (k == 1) {
      carry = (r + a.bc1) * T.sc(0);
    } else {
      lsolve_step(T, k - 2, r * T.sc(k - 1), carry, yprev);
      y[o - P] = yprev;
    }
  }
  y[(long long)(n - 1) * P + l] = carry;
}

// ImplicitAndUpdateQR.F90:39-63 + the forward half of SolveImpEqnUpdate_QR:
// nr rows with identity rows (zero right-hand side) at both walls.
__global__ __launch_bounds__(256) void k_implicit_qr(GridDev g, ImpArgs a,
                                                     const double *__restrict__ qr,
                                                     const double *__restrict__ pr,
                                                     const double *__restrict__ dph,
                                                     double *__restrict__ ru2,
                                                     double *__restrict__ y,
                                                     const double *__restrict__ lu, int nt) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.P) return;
  const int n = g.nr;
  const LUView T = lu_view(lu, nt);
  const long long P = g.P;
  double qm, q0 = qr[l], qp = qr[l + P];
  double carry = 0.0, yprev;  // row 0: zero right-hand side
  for (int k = 2; k <= g.nrm; k++) {
    const long long o = (long long)(k - 1) * P + l;
    qm = q0;
    q0 = qp;
    qp = qr[o + P];
    const double dv = dph[o];
    const double r = rhs_qr(g.hp[k], a, qm, q0, qp, pr[o], pr[o - P], dv, ru2[o]);
    ru2[o] = dv;
    lsolve_step(T, k - 2, r * T.sc(k - 1), carry, yprev);
    y[o - P] = yprev;
  }
  lsolve_step(T, n - 2, 0.0, carry, yprev);  // row nr: zero right-hand side
  y[(long long)(n - 2) * P + l] = yprev;
  y[(long long)(n - 1) * P + l] = carry;
}

// ImplicitAndUpdateQZ.F90:36-115 + SolveImpEqnUpdate_QZ.F90:32-68.
__global__ __launch_bounds__(256) void k_implicit_qz(GridDev g, ImpArgs a,
                                                     const double *__restrict__ qz,
                                                     const double *__restrict__ pr,
                                                     const double *__restrict__ qcap,
                                                     double *__restrict__ ru3,
                                                     double *__restrict__ y,
                                                     const unsigned char *__restrict__ icyl,
                                                     const unsigned char *__restrict__ ocyl,
                                                     const double *__restrict__ lu, int nt) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.P) return;
  const int j = (int)(l / g.nthm), i = (int)(l - (long long)j * g.nthm);
  const long long ljm = (long long)wrap_m(j, g.nzm) * g.nthm + i;
  const int n = g.nrm;
  const int ic = icyl[l], oc = ocyl[l];
  const double fic = (double)ic, foc = (double)oc;
  const LUView T = lu_view(lu + (size_t)(ic + 2 * oc) * nt, nt);
  const long long P = g.P;
  double qm = 0.0, q0 = qz[l], qp = qz[l + P];
  double carry = 0.0, yprev;
  for (int k = 1; k <= n; k++) {
    const long long o = (long long)(k - 1) * P + l;
    if (k > 1) {
      qm = q0;
      q0 = qp;
      qp = (k < n) ? qz[o + P] : 0.0;
    }
    const double qv = qcap[o];
    const double r = rhs_qz(g, g.hp[k], a, k, qm, q0, qp, pr[o], pr[o - l + ljm], qv, ru3[o], fic, foc);
    ru3[o] = qv;
    if (k == 1) {
      carry = r * T.sc(0);
    } else {
      lsolve_step(T, k - 2, r * T.sc(k - 1), carry, yprev);
      y[o - P] = yprev;
    }
  }
  y[(long long)(n - 1) * P + l] = carry;
}

// Backward substitution (dgtts2) of one line: q += x, rows first..n-1 (the
// q_r wall rows are reset instead).  Loads are issued a chunk of rows ahead
// of the dependent recurrence.
// FOLD: y already holds U q + y (the forward sweep added it), so the
// solution of U x' = y' is x' = q + x itself and q is not read.
template <bool FOLD>
DEVI void usolve_line(const LUView &T, int n, int first, double *__restrict__ q,
                      const double *__restrict__ y, long long l, long long P) {
  constexpr int CH = 16;
  auto put = [&](long long o, double qv, double x) { st_nt(q + o, FOLD ? x : qv + x); };
  double x1 = y[(long long)(n - 1) * P + l] * T.rd(n - 1);
  if (n - 1 >= first) {
    const long long o = (long long)(n - 1) * P + l;
    put(o, FOLD ? 0.0 : q[o], x1);
  }
  double x0 = (y[(long long)(n - 2) * P + l] - T.du(n - 2) * x1) * T.rd(n - 2);
  if (n - 2 >= first) {
    const long long o = (long long)(n - 2) * P + l;
    put(o, FOLD ? 0.0 : q[o], x0);
  }
  for (int r0 = n - 3; r0 >= first; r0 -= CH) {
    double yv[CH], qv[CH];
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int r = r0 - u;
      if (r >= first) {
        const long long o = (long long)r * P + l;
        yv[u] = ld_nt(y + o);
        qv[u] = FOLD ? 0.0 : q[o];
      }
    }
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int r = r0 - u;
      if (r >= first) {
        const double x = (yv[u] - T.du(r) * x0 - T.du2(r) * x1) * T.rd(r);
        put((long long)r * P + l, qv[u], x);
        x1 = x0;
        x0 = x;
      }
    }
  }
}

// which: bit 0 q_theta, bit 1 q_r, bit 2 q_z
template <bool FOLD>
DEVI void helm_bwd_body(const long long l, const GridDev &g, double *__restrict__ qt,
                        double *__restrict__ qr, double *__restrict__ qz,
                        const double *__restrict__ yt, const double *__restrict__ yr,
                        const double *__restrict__ yz, const unsigned char *__restrict__ icyl,
                        const unsigned char *__restrict__ ocyl, const double *__restrict__ lut,
                        const double *__restrict__ lur, const double *__restrict__ luz, int nt,
                        int which) {
  if (l >= g.P) return;
  const int v = icyl[l] + 2 * ocyl[l];
  if (which & 1) usolve_line<FOLD>(lu_view(lut + (size_t)v * nt, nt), g.nrm, 0, qt, yt, l, g.P);
  if (which & 2) {
    // SolveImpEqnUpdate_QR.F90:63-65 adds rows 2..nrm; walls reset
    // (ImplicitAndUpdateQR.F90:70-71)
    usolve_line<FOLD>(lu_view(lur, nt), g.nr, 1, qr, yr, l, g.P);
    qr[l] = 0.0;
    qr[(long long)(g.nr - 1) * g.P + l] = 0.0;
  }
  if (which & 4) usolve_line<FOLD>(lu_view(luz + (size_t)v * nt, nt), g.nrm, 0, qz, yz, l, g.P);
}

template <bool FOLD>
__global__ __launch_bounds__(256) void k_helm_bwd(GridDev g, double *__restrict__ qt,
                                                  double *__restrict__ qr,
                                                  double *__restrict__ qz,
                                                  const double *__restrict__ yt,
                                                  const double *__restrict__ yr,
                                                  const double *__restrict__ yz,
                                                  const unsigned char *__restrict__ icyl,
                                                  const unsigned char *__restrict__ ocyl,
                                                  const double *__restrict__ lut,
                                                  const double *__restrict__ lur,
                                                  const double *__restrict__ luz, int nt,
                                                  int which) {
  helm_bwd_body<FOLD>((long long)blockIdx.x * blockDim.x + threadIdx.x, g, qt, qr, qz, yt, yr, yz,
                      icyl, ocyl, lut, lur, luz, nt, which);
}

// Grouped fused backward sweep: blockIdx.y = instance; y' = rhs, dq, qcap.
__global__ __launch_bounds__(256) void k_helm_bwd_g(const StepDesc *__restrict__ tab) {
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  helm_bwd_body<true>((long long)blockIdx.x * blockDim.x + threadIdx.x, glob(d.g),
                      glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]), glob(d.f[TC_RHS]),
                      glob(d.f[TC_DQ]), glob(d.f[TC_QCAP]), glob(d.icyl), glob(d.ocyl),
                      glob(d.lut), glob(d.lur), glob(d.luz), d.lu_nt, 7);
}

// ---------------------------------------------------------------------------
// Fused substep front: explicit terms + right-hand sides + forward
// substitution of all three components in one march along r.  A block owns
// a 64 (theta) x 4 (z) tile of radial lines; the planes k-1, k, k+1 of
// qt, qr, qz (with a one-point periodic halo) rotate through an LDS ring and
// plane k+2 is fetched into registers while plane k is computed.  The
// explicit terms never reach HBM.
// ---------------------------------------------------------------------------
// One point of CorrectVelocity.F90:32-40 and CorrectPressure.F90:37-46
// (amm = apph), without FMA contraction: the pair kernel and the Helmholtz
// kernel that applies a deferred correction round the same way.
// dm: the dph neighbour at i-1 (qt), k-1 (qr), j-1 (qz).
DEVI double corr_qt(const GridDev &g, int kc, double t, double d, double dm, double dt, double al) {
#pragma clang fp contract(off)
  return t - (d - dm) * dt * al * g.dxt * g.usrm[kc];
}
DEVI double corr_qr(const GridDev &g, int kc, double r, double d, double dm, double dt, double al) {
#pragma clang fp contract(off)
  return r - (d - dm) * dt * al * g.rudx2c[kc];
}
DEVI double corr_qz(const GridDev &g, double z, double d, double dm, double dt, double al) {
#pragma clang fp contract(off)
  return z - (d - dm) * al * dt * g.dxz;
}
DEVI double corr_pr(const GridDev &g, int kc, double p, double d, double dim, double dip, double djm,
                    double djp, double dkm, double dkp, double be) {
#pragma clang fp contract(off)
  const double amm = g.apph[kc], acc = g.acph[kc], app = g.apph[kc];
  return p + d -
         be * ((dip - 2.0 * d + dim) * g.udx1qm[kc] + (djp - 2.0 * d + djm) * g.dxzq +
               (dkp * app + d * acc + dkm * amm));
}

constexpr int FT_I = 64, FT_W = FT_I + 2;

// A pressure correction still to be applied: the Helmholtz kernel of the
// next substep applies it to every plane it stages (velocity and pressure)
// and writes the corrected pressure of its tile to pr_out.  The two arrays
// travel as separate __restrict__ arguments: through a struct member the
// compiler could not tell that the pr_out stores leave the grid tables
// alone, and it then loads those with vector loads whose waits also drain
// the plane prefetch.
struct CorrIn {
  double dt, al, be;  // al of the substep that produced dph
};

// CARRY: plane k-1 is not read from the ring but from values the thread
// kept from the previous plane.  The explicit terms and right-hand sides
// read plane k-1 only at (dj, di) = (0, 0), (0, +1) and (+1, 0): m00, m01
// and m10 hold those (the centre, theta+1 and z+1 values of the previous
// iteration's plane k).
template <bool CARRY>
struct LAcc {
  const double *m, *c, *p;  // planes k-1, k, k+1 of one field
  int base;                 // (row+1)*FT_W + tx+1
  double m00, m01, m10;
  DEVI double operator()(int dk, int dj, int di) const {
    if constexpr (CARRY) {
      if (dk < 0) return dj != 0 ? m10 : (di != 0 ? m01 : m00);
    }
    const double *s = dk < 0 ? m : (dk > 0 ? p : c);
    return s[base + dj * FT_W + di];
  }
};

// Workgroups are dispatched round-robin over the XCDs; hand each XCD a
// contiguous run of tiles so neighbouring tiles share one L2.
DEVI int xcd_tile(int b, int nb) {
  constexpr int NX = 8;
  const int q = nb / NX, r = nb - q * NX;
  const int x = b % NX, local = b / NX;
  return x * q + (x < r ? x : r) + local;
}

// A block of 256 threads owns a 64 (theta) x 4*LPT (z) tile of radial lines,
// LPT lines per thread (rows ty and ty+4).  UNIFORM: every wall mask is 1
// (no-slip everywhere, the SlipBCs default), so all lines share one table
// entry and the table pointers stay scalar.
// CORR: qt, qr, qz hold the previous substep's velocity before its pressure
// correction and pr its pressure before the correction; the correction
// (dph of that substep) is applied to each plane as it is staged, from a
// ring of dph planes with a two-point theta/z halo, so the corrected fields
// are the ones the per-substep path would have stored, bit for bit.
template <bool UNIFORM, int LPT, bool CORR>
DEVI void helm_fwd_body(
    const int tile, const GridDev &g, const ImpArgs &a, const double *__restrict__ qt,
    const double *__restrict__ qr, const double *__restrict__ qz, const double *__restrict__ pr,
    double *__restrict__ ru1, double *__restrict__ ru2, double *__restrict__ ru3,
    double *__restrict__ yt, double *__restrict__ yr, double *__restrict__ yz,
    const unsigned char *__restrict__ icyl, const unsigned char *__restrict__ ocyl,
    const double *__restrict__ lut, const double *__restrict__ lur,
    const double *__restrict__ luz, int nt, int ntx, const CorrIn &ci,
    const double *__restrict__ dph, double *__restrict__ pr_out) {
  constexpr int TJ = 4 * LPT, H = TJ + 2, S = FT_W * H, NE = (S + 255) / 256;
  constexpr int W2 = FT_I + 4, S2 = W2 * (TJ + 4), NE2 = (S2 + 255) / 256;
  __shared__ double sq[3][3][S];  // field, ring slot, tile
  __shared__ double sp[2][S];     // pressure planes k and k+1
  __shared__ double sd[CORR ? 3 : 1][CORR ? S2 : 1];  // dph planes (slot = plane % 3)
  const int i0 = (tile % ntx) * FT_I, j0 = (tile / ntx) * TJ;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const long long P = g.P;
  // the tile elements this thread stages per plane: chunks 0..NE-2 are
  // full; in the last one threads past S load a valid element and do not
  // store it, so every global load is unconditional
  long long off[NE];
  const bool has_last = threadIdx.x + (NE - 1) * 256 < S;
#pragma unroll
  for (int q = 0; q < NE; q++) {
    const int e = threadIdx.x + q * 256;
    const int ee = e < S ? e : 0;
    int jj = j0 + ee / FT_W - 1, ii = i0 + ee % FT_W - 1;
    jj = ((jj % g.nzm) + g.nzm) % g.nzm;
    ii = ((ii % g.nthm) + g.nthm) % g.nthm;
    off[q] = (long long)jj * g.nthm + ii;
  }
  // CORR: the dph tile elements (two-point halo) this thread stages, and
  // whether each of its velocity tile elements is an interior point whose
  // corrected pressure it writes
  int off2[CORR ? NE2 : 1];
  bool wpr[NE];
  const bool has_last2 = threadIdx.x + (NE2 - 1) * 256 < S2;
  if constexpr (CORR) {
#pragma unroll
    for (int q = 0; q < NE2; q++) {
      const int e = threadIdx.x + q * 256;
      const int ee = e < S2 ? e : 0;
      int jj = j0 + ee / W2 - 2, ii = i0 + ee % W2 - 2;
      jj = ((jj % g.nzm) + g.nzm) % g.nzm;
      ii = ((ii % g.nthm) + g.nthm) % g.nthm;
      off2[q] = jj * g.nthm + ii;
    }
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = threadIdx.x + q * 256;
      const int row = e / FT_W, col = e % FT_W;
      wpr[q] = (e < S) && row >= 1 && row <= TJ && col >= 1 && col <= FT_I &&
               j0 + row - 1 < g.nzm && i0 + col - 1 < g.nthm;
    }
  }
  // the corrected velocity of plane p at tile element e (correct_pair's
  // arithmetic; q_r is zero on both walls)
  auto corr_q = [&](int p, int e, double &vt, double &vr, double &vz) {
    const int row = e / FT_W, col = e - row * FT_W;
    const int e2 = (row + 1) * W2 + col + 1;
    const double *dc = sd[CORR ? p % 3 : 0];
    if (p <= g.nrm) {
      vt = corr_qt(g, p, vt, dc[e2], dc[e2 - 1], ci.dt, ci.al);
      vz = corr_qz(g, vz, dc[e2], dc[e2 - W2], ci.dt, ci.al);
    }
    if (p == 1 || p >= g.nr)
      vr = 0.0;
    else
      vr = corr_qr(g, p, vr, dc[e2], sd[CORR ? (p - 1) % 3 : 0][e2], ci.dt, ci.al);
  };
  // the corrected pressure of plane p <= nrm at tile element e
  auto corr_p = [&](int p, int e, double v) {
    const int row = e / FT_W, col = e - row * FT_W;
    const int e2 = (row + 1) * W2 + col + 1;
    const double *dc = sd[CORR ? p % 3 : 0];
    const double *dkp = sd[CORR ? (p < g.nrm ? p + 1 : p) % 3 : 0];
    const double *dkm = sd[CORR ? (p > 1 ? p - 1 : p) % 3 : 0];
    return corr_pr(g, p, v, dc[e2], dc[e2 - 1], dc[e2 + 1], dc[e2 - W2], dc[e2 + W2], dkm[e2],
                   dkp[e2], ci.be);
  };
  // the RK history ru enters as ro*ru only: with ro = 0 (the first RK3
  // substep) it is not read at all
  const bool hist = CORR || a.ro != 0.0;
  // per-line state
  bool valid[LPT];
  long long l[LPT];
  int base[LPT];
  double fic[LPT], foc[LPT], h1[LPT], h2[LPT], h3[LPT], ct[LPT], cr[LPT], cz[LPT], pr_km[LPT];
  // !CORR: plane k-1 values carried from the previous iteration (qt centre,
  // qt theta+1, qr centre, qz centre, qz z+1), so that no thread reads the
  // ring slot of plane k-1 and the staging of plane k+2 into it needs no
  // barrier of its own
  double cm[LPT][5];
  LUView Tt[LPT], Tz[LPT];
  const LUView Tr = lu_view(lur, nt);
#pragma unroll
  for (int u = 0; u < LPT; u++) {
    const int row = ty + 4 * u;
    const int i = i0 + tx, j = j0 + row;
    valid[u] = (i < g.nthm) && (j < g.nzm);
    l[u] = valid[u] ? (long long)j * g.nthm + i : 0;
    base[u] = (row + 1) * FT_W + tx + 1;
    const int ic = (UNIFORM || !valid[u]) ? 1 : icyl[l[u]];
    const int oc = (UNIFORM || !valid[u]) ? 1 : ocyl[l[u]];
    fic[u] = (double)ic;
    foc[u] = (double)oc;
    const int v = UNIFORM ? 3 : ic + 2 * oc;
    Tt[u] = lu_view(lut + (size_t)v * nt, nt);
    Tz[u] = lu_view(luz + (size_t)v * nt, nt);
    h1[u] = (valid[u] && hist) ? ld_nt(ru1 + l[u]) : 0.0;
    h2[u] = 0.0;
    h3[u] = (valid[u] && hist) ? ld_nt(ru3 + l[u]) : 0.0;
    ct[u] = cr[u] = cz[u] = pr_km[u] = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) cm[u][q] = 0.0;
  }
  // planes 1 and 2 of the velocity, plane 1 of the pressure (CORR: planes 1
  // and 2 of dph first, then the corrected values)
  if constexpr (CORR) {
    const long long o2 = (g.nrm > 1) ? P : 0;
#pragma unroll
    for (int q = 0; q < NE2; q++) {
      const double a1 = dph[off2[q]], a2 = dph[o2 + off2[q]];
      if (q < NE2 - 1 || has_last2) {
        const int e = threadIdx.x + q * 256;
        sd[1][e] = a1;
        sd[2][e] = a2;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NE; q++) {
    double v0 = qt[off[q]], v1 = qr[off[q]], v2 = qz[off[q]];
    double w0 = qt[P + off[q]], w1 = qr[P + off[q]], w2 = qz[P + off[q]];
    double v3 = pr[off[q]];
    if (q < NE - 1 || has_last) {
      if constexpr (CORR) {
        const int e = threadIdx.x + q * 256;
        corr_q(1, e, v0, v1, v2);
        corr_q(2, e, w0, w1, w2);
        v3 = corr_p(1, e, v3);
        if (wpr[q]) st_nt(pr_out + off[q], v3);
      }
      const int e = threadIdx.x + q * 256;
      sq[0][1][e] = v0;
      sq[1][1][e] = v1;
      sq[2][1][e] = v2;
      sq[0][2][e] = w0;
      sq[1][2][e] = w1;
      sq[2][2][e] = w2;
      sp[1][e] = v3;
    }
  }
  __syncthreads();
  const int nrm = g.nrm;
  double yv;

  for (int k = 1; k <= nrm; k++) {
    // prefetch plane k+2 of the velocity, plane k+1 of the pressure and of
    // the RK histories.  In the last iteration these re-read the last plane
    // (their values are not used), so the loads carry no condition.
    double nq[3][NE], np[NE], n1[LPT], n2[LPT], n3[LPT], nd[CORR ? NE2 : 1];
    const long long oq = (long long)(k + 2 <= g.nr ? k + 1 : g.nr - 1) * P;
    const long long op = (long long)(k + 1 <= nrm ? k : nrm - 1) * P;
#pragma unroll
    for (int q = 0; q < NE; q++) {
      nq[0][q] = qt[oq + off[q]];
      nq[1][q] = qr[oq + off[q]];
      nq[2][q] = qz[oq + off[q]];
      np[q] = pr[op + off[q]];
    }
    if constexpr (CORR) {
      // dph plane k+2 (clamped to nrm: beyond it no correction reads dph)
      const long long od = (long long)(k + 2 <= nrm ? k + 1 : nrm - 1) * P;
#pragma unroll
      for (int q = 0; q < NE2; q++) nd[q] = dph[od + off2[q]];
    }
    if (hist) {
#pragma unroll
      for (int u = 0; u < LPT; u++) {
        n1[u] = ld_nt(ru1 + op + l[u]);
        n2[u] = ld_nt(ru2 + op + l[u]);
        n3[u] = ld_nt(ru3 + op + l[u]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < LPT; u++) n1[u] = n2[u] = n3[u] = 0.0;
    }
    const int sm = (k + 2) % 3, sc = k % 3, sn = (k + 1) % 3;  // planes k-1, k, k+1
    const double *spk = sp[k & 1];
    const HelmPlane hc = g.hp[k];
#pragma unroll
    for (int u = 0; u < LPT; u++) {
      if (!valid[u]) continue;
      LAcc<!CORR> T, R, Z;
      T.m = sq[0][sm]; T.c = sq[0][sc]; T.p = sq[0][sn]; T.base = base[u];
      R.m = sq[1][sm]; R.c = sq[1][sc]; R.p = sq[1][sn]; R.base = base[u];
      Z.m = sq[2][sm]; Z.c = sq[2][sc]; Z.p = sq[2][sn]; Z.base = base[u];
      T.m00 = cm[u][0]; T.m01 = cm[u][1]; T.m10 = 0.0;
      R.m00 = cm[u][2]; R.m01 = R.m10 = 0.0;
      Z.m00 = cm[u][3]; Z.m01 = 0.0; Z.m10 = cm[u][4];
      const double pr_c = spk[base[u]], pr_im = spk[base[u] - 1], pr_jm = spk[base[u] - FT_W];
      const long long o = (long long)(k - 1) * P + l[u];
      // q_theta
      {
        const double dqv = explicit_qt_g(g, hc, T, R, Z, k);
        const double qm = (k > 1) ? T(-1, 0, 0) : 0.0, qp = (k < nrm) ? T(1, 0, 0) : 0.0;
        const double r = rhs_qt<UNIFORM>(g, hc, a, k, qm, T(0, 0, 0), qp, pr_c, pr_im, dqv, h1[u], fic[u], foc[u]);
        st_nt(ru1 + o, dqv);
        if (k == 1) {
          ct[u] = (r + a.bc1) * Tt[u].sc(0);
        } else {
          lsolve_step(Tt[u], k - 2, r * Tt[u].sc(k - 1), ct[u], yv);
          st_nt(yt + o - P, yv + Tt[u].uq(k - 2, qm, T(0, 0, 0), qp));
        }
      }
      // q_r (rows 2..nrm; row 1 is the wall with zero right-hand side)
      if (k >= 2) {
        const double dphv = explicit_qr_g(g, hc, T, R, Z, k);
        const double r = rhs_qr(hc, a, R(-1, 0, 0), R(0, 0, 0), R(1, 0, 0), pr_c, pr_km[u], dphv, h2[u]);
        st_nt(ru2 + o, dphv);
        lsolve_step(Tr, k - 2, r * Tr.sc(k - 1), cr[u], yv);
        st_nt(yr + o - P, yv + Tr.uq(k - 2, R(-1, 0, 0), R(0, 0, 0), R(1, 0, 0)));
      }
      // q_z
      {
        const double qcv = explicit_qz_g(g, hc, T, R, Z, k);
        const double qm = (k > 1) ? Z(-1, 0, 0) : 0.0, qp = (k < nrm) ? Z(1, 0, 0) : 0.0;
        const double r = rhs_qz<UNIFORM>(g, hc, a, k, qm, Z(0, 0, 0), qp, pr_c, pr_jm, qcv, h3[u], fic[u], foc[u]);
        st_nt(ru3 + o, qcv);
        if (k == 1) {
          cz[u] = r * Tz[u].sc(0);
        } else {
          lsolve_step(Tz[u], k - 2, r * Tz[u].sc(k - 1), cz[u], yv);
          st_nt(yz + o - P, yv + Tz[u].uq(k - 2, qm, Z(0, 0, 0), qp));
        }
      }
      pr_km[u] = pr_c;
      if constexpr (!CORR) {
        cm[u][0] = T(0, 0, 0);
        cm[u][1] = T(0, 0, 1);
        cm[u][2] = R(0, 0, 0);
        cm[u][3] = Z(0, 0, 0);
        cm[u][4] = Z(0, 1, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < LPT; u++) {
      h1[u] = n1[u];
      h2[u] = n2[u];
      h3[u] = n3[u];
    }
    if constexpr (CORR) {
      // dph slot (k+2)%3 held plane k-1, last read before the previous barrier
#pragma unroll
      for (int q = 0; q < NE2; q++)
        if (q < NE2 - 1 || has_last2) sd[(k + 2) % 3][threadIdx.x + q * 256] = nd[q];
      // the ring slot of plane k-1 is read above, and the dph plane k+2 just
      // written is read by the staging below
      __syncthreads();
    }
    // in the last iteration these fill slots no later plane reads
    const int s2 = (k + 2) % 3;
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = threadIdx.x + q * 256;
      if (q < NE - 1 || has_last) {
        if constexpr (CORR) {
          corr_q(k + 2, e, nq[0][q], nq[1][q], nq[2][q]);
          if (k + 1 <= nrm) {
            np[q] = corr_p(k + 1, e, np[q]);
            if (wpr[q]) st_nt(pr_out + (long long)k * P + off[q], np[q]);
          }
        }
        sq[0][s2][e] = nq[0][q];
        sq[1][s2][e] = nq[1][q];
        sq[2][s2][e] = nq[2][q];
        sp[(k + 1) & 1][e] = np[q];
      }
    }
    __syncthreads();
  }
  // after the last barrier the ring still holds planes nrm and nrm+1
  const int s0 = nrm % 3, s1 = (nrm + 1) % 3;
#pragma unroll
  for (int u = 0; u < LPT; u++) {
    if (!valid[u]) continue;
    const int b = base[u];
    yt[(long long)(nrm - 1) * P + l[u]] = ct[u] + Tt[u].ud(nrm - 1) * sq[0][s0][b];
    yz[(long long)(nrm - 1) * P + l[u]] = cz[u] + Tz[u].ud(nrm - 1) * sq[2][s0][b];
    double c2 = cr[u];
    lsolve_step(Tr, g.nr - 2, 0.0, c2, yv);  // wall row nr
    yr[(long long)(g.nr - 2) * P + l[u]] = yv + Tr.ud(g.nr - 2) * sq[1][s0][b] + Tr.du(g.nr - 2) * sq[1][s1][b];
    yr[(long long)(g.nr - 1) * P + l[u]] = c2 + Tr.ud(g.nr - 1) * sq[1][s1][b];
  }
}

template <bool UNIFORM, int LPT, bool CORR>
__global__ __launch_bounds__(256) void k_helm_fwd(
    GridDev g, ImpArgs a, const double *__restrict__ qt, const double *__restrict__ qr,
    const double *__restrict__ qz, const double *__restrict__ pr, double *__restrict__ ru1,
    double *__restrict__ ru2, double *__restrict__ ru3, double *__restrict__ yt,
    double *__restrict__ yr, double *__restrict__ yz, const unsigned char *__restrict__ icyl,
    const unsigned char *__restrict__ ocyl, const double *__restrict__ lut,
    const double *__restrict__ lur, const double *__restrict__ luz, int nt, int ntx, int ntz,
    CorrIn ci, const double *__restrict__ dph, double *__restrict__ pr_out) {
  helm_fwd_body<UNIFORM, LPT, CORR>(xcd_tile(blockIdx.x, ntx * ntz), g, a, qt, qr, qz, pr, ru1,
                                    ru2, ru3, yt, yr, yz, icyl, ocyl, lut, lur, luz, nt, ntx, ci,
                                    dph, pr_out);
}

// Grouped: blockIdx.x = tile, blockIdx.y = instance of the shape group; the
// XCD remap runs over the flattened (instance, tile) index.
template <bool UNIFORM, int LPT, bool CORR>
__global__ __launch_bounds__(256) void k_helm_fwd_g(const StepDesc *__restrict__ tab,
                                                    const BatchDyn *__restrict__ dyn, int ns,
                                                    int ntx, int ntz) {
  const int ntiles = ntx * ntz;
  const int lin = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int inst = lin / ntiles, tile = lin - inst * ntiles;
  const StepDesc &d = tab[inst];
  if (inst_stopped(d)) return;
  ImpArgs a = d.a;
  a.vwall_in = glob(dyn)[inst].vwall[ns - 1];
  a.bc1 = glob(dyn)[inst].bc1[ns - 1];
  const CorrIn ci{d.dt, d.corr_al, d.beta};
  helm_fwd_body<UNIFORM, LPT, CORR>(
      tile, glob(d.g), a, glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]), glob(d.pr_in),
      glob(d.f[TC_RU1]), glob(d.f[TC_RU2]), glob(d.f[TC_RU3]), glob(d.f[TC_RHS]), glob(d.f[TC_DQ]),
      glob(d.f[TC_QCAP]), glob(d.icyl), glob(d.ocyl), glob(d.lut), glob(d.lur), glob(d.luz),
      d.lu_nt, ntx, ci, glob(d.f[TC_DPH]), glob(d.pr_out));
}

static ImpArgs imp_args(const tc_ctx *c) {
  ImpArgs a;
  a.al = c->al;
  a.ga = c->ga;
  a.ro = c->ro;
  a.dt = c->dt;
  a.alre = c->al / c->p.ren;
  a.vwall_in = c->p.virotint + c->vinner_t;
  a.virotext = c->p.virotext;
  a.bc1 = 0.5 * c->al * c->dt * c->dvinner_t * c->g.am1j[1] / c->p.ren;
  a.udxz = c->g.dxz * c->al;
  return a;
}

static const double *lu_slot(const tc_ctx *c, int comp) {
  // tables laid out [slot = ns-1][comp][variant 0..3][nt]
  return c->lu + ((size_t)(c->ns - 1) * 3 + comp) * 4 * (size_t)c->lu_nt;
}

static unsigned line_blocks(const tc_ctx *c) { return (unsigned)((c->P + 255) / 256); }

static void launch_bwd(tc_ctx *c, int which, const char *name) {
  timer_begin(c, name);
  hipLaunchKernelGGL(k_helm_bwd<true>, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_RHS], c->f[TC_DQ],
                     c->f[TC_QCAP], c->mask, c->mask + c->P, lu_slot(c, LU_QT), lu_slot(c, LU_QR),
                     lu_slot(c, LU_QZ), c->lu_nt, which);
  timer_end(c);
}

// Routine-level entries: right-hand side + forward substitution into the rhs
// field, then the backward substitution.
void launch_implicit_qt(tc_ctx *c) {
  timer_begin(c, "implicit_qt");
  hipLaunchKernelGGL(k_implicit_qt, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QT], c->f[TC_PR], c->f[TC_DQ], c->f[TC_RU1],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QT), c->lu_nt);
  timer_end(c);
  hipLaunchKernelGGL(k_helm_bwd<false>, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_RHS], c->f[TC_RHS],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QT), lu_slot(c, LU_QR),
                     lu_slot(c, LU_QZ), c->lu_nt, 1);
}
void launch_implicit_qr(tc_ctx *c) {
  timer_begin(c, "implicit_qr");
  hipLaunchKernelGGL(k_implicit_qr, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QR], c->f[TC_PR], c->f[TC_DPH], c->f[TC_RU2],
                     c->f[TC_RHS], lu_slot(c, LU_QR), c->lu_nt);
  timer_end(c);
  hipLaunchKernelGGL(k_helm_bwd<false>, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_RHS], c->f[TC_RHS],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QT), lu_slot(c, LU_QR),
                     lu_slot(c, LU_QZ), c->lu_nt, 2);
}
void launch_implicit_qz(tc_ctx *c) {
  timer_begin(c, "implicit_qz");
  hipLaunchKernelGGL(k_implicit_qz, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QZ], c->f[TC_PR], c->f[TC_QCAP], c->f[TC_RU3],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QZ), c->lu_nt);
  timer_end(c);
  hipLaunchKernelGGL(k_helm_bwd<false>, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_RHS], c->f[TC_RHS],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QT), lu_slot(c, LU_QR),
                     lu_slot(c, LU_QZ), c->lu_nt, 4);
}

// Fused path: y_theta -> rhs, y_r -> dq, y_z -> qcap (the explicit terms
// stay in registers, so those fields serve as scratch).
template <bool U, int LPT, bool CORR>
static void launch_fwd(tc_ctx *c, const double *pr, const CorrIn &ci, double *pr_out) {
  const int ntx = (c->nthm + FT_I - 1) / FT_I, ntz = (c->nzm + 4 * LPT - 1) / (4 * LPT);
  hipLaunchKernelGGL((k_helm_fwd<U, LPT, CORR>), dim3((unsigned)(ntx * ntz)), dim3(256), 0,
                     c->stream, c->gd, imp_args(c), c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], pr,
                     c->f[TC_RU1], c->f[TC_RU2], c->f[TC_RU3], c->f[TC_RHS], c->f[TC_DQ],
                     c->f[TC_QCAP], c->mask, c->mask + c->P, lu_slot(c, LU_QT), lu_slot(c, LU_QR),
                     lu_slot(c, LU_QZ), c->lu_nt, ntx, ntz, ci, c->f[TC_DPH], pr_out);
}

// Lines per thread of the fused kernel: two where the two-line tiles alone
// fill every resident slot of the device at least once, one otherwise.
// Two lines per thread give each block twice the work per plane for less
// than twice the time, but half the tiles: at 512x512x385 the 512 two-line
// tiles fill the 512 slots once (2771 us per launch against 3445 for one
// line), at 1024x1024x769 four times (22.4 ms against 26.0), while a lone
// 256^2x193 grid (128 two-line tiles) runs faster with one line.
int choose_helm_lpt(const tc_ctx *c) {
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) return 1;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_helm_fwd<true, 2, true>, 256, 0) !=
          hipSuccess ||
      per_cu <= 0)
    return 1;
  const long long tiles = (long long)((c->nthm + FT_I - 1) / FT_I) * ((c->nzm + 7) / 8);
  return tiles >= (long long)per_cu * cus ? 2 : 1;
}

void launch_helmholtz_fused(tc_ctx *c) {
  timer_begin(c, "helm_fwd");
  const CorrIn ci{};
  if (c->masks_uniform_one) {
    if (c->helm_lpt == 2) launch_fwd<true, 2, false>(c, c->f[TC_PR], ci, nullptr);
    else launch_fwd<true, 1, false>(c, c->f[TC_PR], ci, nullptr);
  } else {
    if (c->helm_lpt == 2) launch_fwd<false, 2, false>(c, c->f[TC_PR], ci, nullptr);
    else launch_fwd<false, 1, false>(c, c->f[TC_PR], ci, nullptr);
  }
  timer_end(c);
  launch_bwd(c, 7, "helm_bwd");
}

// The same with the previous substep's pressure correction (dph, al_prev)
// applied while staging: the velocity fields hold that substep's uncorrected
// velocity, pr_in its uncorrected pressure; the corrected pressure goes to
// pr_out (a different array: neighbouring tiles still read pr_in).
void launch_helmholtz_fused_corr(tc_ctx *c, const double *pr_in, double *pr_out, double al_prev) {
  timer_begin(c, "helm_fwd_corr");
  const CorrIn ci{c->dt, al_prev, c->beta};
  if (c->masks_uniform_one) {
    if (c->helm_lpt == 2) launch_fwd<true, 2, true>(c, pr_in, ci, pr_out);
    else launch_fwd<true, 1, true>(c, pr_in, ci, pr_out);
  } else {
    if (c->helm_lpt == 2) launch_fwd<false, 2, true>(c, pr_in, ci, pr_out);
    else launch_fwd<false, 1, true>(c, pr_in, ci, pr_out);
  }
  timer_end(c);
  launch_bwd(c, 7, "helm_bwd");
}

// ---------------------------------------------------------------------------
// Divergence and projection updates, one point per thread
// ---------------------------------------------------------------------------

DEVI double divergence_at(const GridDev &g, const double *__restrict__ qt,
                          const double *__restrict__ qr, const double *__restrict__ qz, int kc,
                          int j, int i) {
  // CalcLocalDivergence.F90:33-35 (also CheckDivergence.F90:34-36)
  const int ip = wrap_p(i, g.nthm), jp = wrap_p(j, g.nzm);
  const long long o = IX(g, kc, j, i);
  return ((qt[IX(g, kc, j, ip)] - qt[o]) * g.dxt + (qr[o + g.P] - qr[o]) * g.udx2m[kc]) * g.usrm[kc] +
         (qz[IX(g, kc, jp, i)] - qz[o]) * g.dxz;
}

__global__ __launch_bounds__(256) void k_divergence(GridDev g, const double *__restrict__ qt,
                                                    const double *__restrict__ qr,
                                                    const double *__restrict__ qz,
                                                    double *__restrict__ dph, double usdt) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  const int kc = blockIdx.z + 1;
  dph[IX(g, kc, j, i)] = divergence_at(g, qt, qr, qz, kc, j, i) * usdt;
}

void launch_divergence(tc_ctx *c, double usdt) {
  timer_begin(c, "divergence");
  hipLaunchKernelGGL(k_divergence, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_DPH], usdt);
  timer_end(c);
}

// CorrectVelocity.F90:32-40 (qr walls zeroed by k_zero_wall_qr)
DEVI void correct_velocity_at(const GridDev &g, double *__restrict__ qt, double *__restrict__ qr,
                              double *__restrict__ qz, const double *__restrict__ dph, int kc,
                              int j, int i, double dt, double al) {
  const int im = wrap_m(i, g.nthm), jm = wrap_m(j, g.nzm);
  const int km = kc > 1 ? kc - 1 : 1;
  const long long o = IX(g, kc, j, i);
  const double locdph = dph[o];
  qt[o] = qt[o] - (locdph - dph[IX(g, kc, j, im)]) * dt * al * g.dxt * g.usrm[kc];
  qr[o] = qr[o] - (locdph - dph[IX(g, km, j, i)]) * dt * al * g.rudx2c[kc];
  qz[o] = qz[o] - (locdph - dph[IX(g, kc, jm, i)]) * al * dt * g.dxz;
}

// CorrectPressure.F90:37-46 (amm = apph, be = beta)
DEVI void correct_pressure_at(const GridDev &g, double *__restrict__ pr,
                              const double *__restrict__ dph, int kc, int j, int i, double be) {
  const int ip = wrap_p(i, g.nthm), im = wrap_m(i, g.nthm);
  const int jp = wrap_p(j, g.nzm), jm = wrap_m(j, g.nzm);
  const int kp = kc < g.nrm ? kc + 1 : g.nrm, km = kc > 1 ? kc - 1 : 1;
  const double amm = g.apph[kc], acc = g.acph[kc], app = g.apph[kc];
  const long long o = IX(g, kc, j, i);
  const double d0 = dph[o];
  pr[o] = pr[o] + d0 -
          be * ((dph[IX(g, kc, j, ip)] - 2.0 * d0 + dph[IX(g, kc, j, im)]) * g.udx1qm[kc] +
                (dph[IX(g, kc, jp, i)] - 2.0 * d0 + dph[IX(g, kc, jm, i)]) * g.dxzq +
                (dph[IX(g, kp, j, i)] * app + d0 * acc + dph[IX(g, km, j, i)] * amm));
}

__global__ __launch_bounds__(256) void k_correct_velocity(GridDev g, double *__restrict__ qt,
                                                          double *__restrict__ qr,
                                                          double *__restrict__ qz,
                                                          const double *__restrict__ dph,
                                                          double dt, double al) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  correct_velocity_at(g, qt, qr, qz, dph, blockIdx.z + 1, j, i, dt, al);
}

__global__ __launch_bounds__(256) void k_correct_pressure(GridDev g, double *__restrict__ pr,
                                                          const double *__restrict__ dph,
                                                          double be) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  correct_pressure_at(g, pr, dph, blockIdx.z + 1, j, i, be);
}

__global__ __launch_bounds__(256) void k_correct_both(GridDev g, double *__restrict__ qt,
                                                      double *__restrict__ qr,
                                                      double *__restrict__ qz,
                                                      double *__restrict__ pr,
                                                      const double *__restrict__ dph, double dt,
                                                      double al, double be) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  const int kc = blockIdx.z + 1;
  correct_velocity_at(g, qt, qr, qz, dph, kc, j, i, dt, al);
  correct_pressure_at(g, pr, dph, kc, j, i, be);
  if (kc == 1) qr[IX(g, 1, j, i)] = 0.0;
  if (kc == g.nrm) qr[IX(g, g.nr, j, i)] = 0.0;
}

// CorrectVelocity + CorrectPressure + the qr wall reset on a theta pair
// (nthm even): one lane owns points i, i+1 (i even) and moves them as 16-byte
// loads and stores; the arithmetic per point is that of correct_velocity_at
// and correct_pressure_at.
DEVI void correct_pair(const GridDev &g, double *__restrict__ qt, double *__restrict__ qr,
                       double *__restrict__ qz, double *__restrict__ pr,
                       const double *__restrict__ dph, int kc, int j, int i, double dt, double al,
                       double be) {
  // no FMA contraction: every caller (plain and grouped) rounds the same way
#pragma clang fp contract(off)
  const int im = wrap_m(i, g.nthm), ip = wrap_p(i + 1, g.nthm);
  const int jp = wrap_p(j, g.nzm), jm = wrap_m(j, g.nzm);
  const int kp = kc < g.nrm ? kc + 1 : g.nrm, km = kc > 1 ? kc - 1 : 1;
  const long long o = IX(g, kc, j, i);
  auto ld2 = [](const double *a, long long x) { return *reinterpret_cast<const double2 *>(a + x); };
  auto st2 = [](double *a, long long x, double u, double v) {
    *reinterpret_cast<double2 *>(a + x) = make_double2(u, v);
  };
  const double2 d = ld2(dph, o), djm = ld2(dph, IX(g, kc, jm, i)), djp = ld2(dph, IX(g, kc, jp, i));
  const double2 dkm = ld2(dph, IX(g, km, j, i)), dkp = ld2(dph, IX(g, kp, j, i));
  const double dim_ = dph[IX(g, kc, j, im)], dip = dph[IX(g, kc, j, ip)];
  const double2 t = ld2_nt(qt + o), r = ld2_nt(qr + o), z = ld2_nt(qz + o), p = ld2_nt(pr + o);
  // CorrectVelocity.F90:32-40
  st2_nt(qt + o, corr_qt(g, kc, t.x, d.x, dim_, dt, al), corr_qt(g, kc, t.y, d.y, d.x, dt, al));
  st2_nt(qr + o, corr_qr(g, kc, r.x, d.x, dkm.x, dt, al), corr_qr(g, kc, r.y, d.y, dkm.y, dt, al));
  st2_nt(qz + o, corr_qz(g, z.x, d.x, djm.x, dt, al), corr_qz(g, z.y, d.y, djm.y, dt, al));
  // CorrectPressure.F90:37-46 (be = beta)
  st2_nt(pr + o, corr_pr(g, kc, p.x, d.x, dim_, d.y, djm.x, djp.x, dkm.x, dkp.x, be),
      corr_pr(g, kc, p.y, d.y, d.x, dip, djm.y, djp.y, dkm.y, dkp.y, be));
  // CorrectVelocity.F90:46-47
  if (kc == 1) st2(qr, IX(g, 1, j, i), 0.0, 0.0);
  if (kc == g.nrm) st2(qr, IX(g, g.nr, j, i), 0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_correct_both2(GridDev g, double *__restrict__ qt,
                                                       double *__restrict__ qr,
                                                       double *__restrict__ qz,
                                                       double *__restrict__ pr,
                                                       const double *__restrict__ dph, double dt,
                                                       double al, double be) {
  const int i = 2 * (blockIdx.x * 64 + threadIdx.x);
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  correct_pair(g, qt, qr, qz, pr, dph, blockIdx.z + 1, j, i, dt, al, be);
}

// Grouped, theta pairs: blockIdx.z = instance * nrm + plane.
__global__ __launch_bounds__(256) void k_correct_both2_g(const StepDesc *__restrict__ tab, int nrm) {
  const int inst = blockIdx.z / nrm, kc = blockIdx.z - inst * nrm + 1;
  const StepDesc &d = tab[inst];
  if (inst_stopped(d)) return;
  const int i = 2 * (blockIdx.x * 64 + threadIdx.x);
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.g.nthm || j >= d.g.nzm) return;
  correct_pair(glob(d.g), glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]), glob(d.f[TC_PR]),
               glob(d.f[TC_DPH]), kc, j, i, d.dt, d.al, d.beta);
}

// Grouped: blockIdx.z = instance * nrm + plane.
__global__ __launch_bounds__(256) void k_correct_both_g(const StepDesc *__restrict__ tab, int nrm) {
  const int inst = blockIdx.z / nrm, kc = blockIdx.z - inst * nrm + 1;
  const StepDesc &d = tab[inst];
  if (inst_stopped(d)) return;
  const GridDev g = glob(d.g);
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  double *qr = glob(d.f[TC_QR]);
  correct_velocity_at(g, glob(d.f[TC_QT]), qr, glob(d.f[TC_QZ]), glob(d.f[TC_DPH]), kc, j, i, d.dt,
                      d.al);
  correct_pressure_at(g, glob(d.f[TC_PR]), glob(d.f[TC_DPH]), kc, j, i, d.beta);
  if (kc == 1) qr[IX(g, 1, j, i)] = 0.0;
  if (kc == g.nrm) qr[IX(g, g.nr, j, i)] = 0.0;
}

// qr(1,:,:) = qr(nr,:,:) = 0 (CorrectVelocity.F90:46-47)
__global__ void k_zero_wall_qr(GridDev g, double *__restrict__ qr) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.P) return;
  qr[l] = 0.0;
  qr[(long long)(g.nr - 1) * g.P + l] = 0.0;
}

void launch_correct_velocity(tc_ctx *c, double aldt_al) {
  timer_begin(c, "correct_velocity");
  hipLaunchKernelGGL(k_correct_velocity, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_DPH], c->dt, aldt_al);
  timer_end(c);
  launch_zero_wall_qr(c);
}
void launch_correct_pressure(tc_ctx *c) {
  timer_begin(c, "correct_pressure");
  hipLaunchKernelGGL(k_correct_pressure, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_PR], c->f[TC_DPH], c->beta);
  timer_end(c);
}
void launch_correct_both(tc_ctx *c) {
  timer_begin(c, "correct_both");
  if (c->nthm % 2 == 0)
    hipLaunchKernelGGL(k_correct_both2,
                       dim3((unsigned)((c->nthm / 2 + 63) / 64), (unsigned)((c->nzm + 3) / 4),
                            (unsigned)c->nrm),
                       dim3(64, 4), 0, c->stream, c->gd, c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ],
                       c->f[TC_PR], c->f[TC_DPH], c->dt, c->al, c->beta);
  else
    hipLaunchKernelGGL(k_correct_both, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                       c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_PR], c->f[TC_DPH], c->dt,
                       c->al, c->beta);
  timer_end(c);
}
void launch_zero_wall_qr(tc_ctx *c) {
  hipLaunchKernelGGL(k_zero_wall_qr, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     c->f[TC_QR]);
}

// ---------------------------------------------------------------------------
// Reductions: per-block partials, then one block per value combines them in
// a fixed order, so sums are deterministic run to run.
// ---------------------------------------------------------------------------

constexpr int RED_THREADS = 256;

template <int NV>
DEVI void block_reduce(double (&v)[NV], const int *ops /*0 sum, 1 max*/, double *out) {
  __shared__ double sh[NV][RED_THREADS / 64];
  // linear thread index: the blocks are 256 x 1 or 64 x 4
  const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
  const int lane = t & 63, wid = t >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_xor(x, off, 64);
      x = ops[q] ? fmax(x, y) : x + y;
    }
    if (lane == 0) sh[q][wid] = x;
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double x = sh[q][0];
      for (int w = 1; w < RED_THREADS / 64; w++) x = ops[q] ? fmax(x, sh[q][w]) : x + sh[q][w];
      out[q] = x;
    }
  }
}

// Point reductions march along r: thread = (j, i) column, k = 1..nrm, so the
// qr(k+1) plane of one step is the qr(k) plane of the next.
struct Col {
  long long l;
  long long lip, ljp;  // in-plane offsets of (j, i+1) and (j+1, i)
  bool ok;
};
DEVI Col column(const GridDev &g, long long blk) {
  Col c;
  c.l = blk * blockDim.x + threadIdx.x;
  c.ok = c.l < g.P;
  const long long l = c.ok ? c.l : 0;
  const int j = (int)(l / g.nthm), i = (int)(l - (long long)j * g.nthm);
  c.lip = (long long)j * g.nthm + wrap_p(i, g.nthm);
  c.ljp = (long long)wrap_p(j, g.nzm) * g.nthm + i;
  return c;
}

// CalcMaxCFL.F90:30-43 (WHICH 0) and CheckDivergence.F90:28-40 (WHICH 1)
template <int WHICH>
DEVI void pointmax_body(const int blk, const GridDev &g, const double *__restrict__ qt,
                        const double *__restrict__ qr, const double *__restrict__ qz,
                        double *__restrict__ part) {
#pragma clang fp contract(off)
  double m[1] = {0.0};
  const Col c = column(g, blk);
  if (c.ok) {
    double r_c = qr[c.l];
    for (int kc = 1; kc <= g.nrm; kc++) {
      const long long b = (long long)(kc - 1) * g.P;
      const double t_c = qt[b + c.l], t_ip = qt[b + c.lip];
      const double r_kp = qr[b + g.P + c.l];
      const double z_c = qz[b + c.l], z_jp = qz[b + c.ljp];
      double val;
      if (WHICH == 0)
        val = (fabs((t_ip + t_c) * 0.5 * g.dxt * g.usrm[kc]) + fabs((r_kp + r_c) * 0.5 * g.udx2rm[kc]) +
               fabs((z_jp + z_c) * 0.5 * g.dxz));
      else
        val = fabs(((t_ip - t_c) * g.dxt + (r_kp - r_c) * g.udx2m[kc]) * g.usrm[kc] + (z_jp - z_c) * g.dxz);
      m[0] = fmax(m[0], val);
      r_c = r_kp;
    }
  }
  const int ops[1] = {1};
  block_reduce<1>(m, ops, part + blk);
}

template <int WHICH>
__global__ __launch_bounds__(RED_THREADS) void k_pointmax(GridDev g, const double *__restrict__ qt,
                                                         const double *__restrict__ qr,
                                                         const double *__restrict__ qz,
                                                         double *__restrict__ part) {
  pointmax_body<WHICH>(blockIdx.x, g, qt, qr, qz, part);
}

// Grouped CalcMaxCFL partials: blockIdx.y = instance.
__global__ __launch_bounds__(RED_THREADS) void k_pointmax_cfl_g(const StepDesc *__restrict__ tab) {
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  pointmax_body<0>(blockIdx.x, glob(d.g), glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]),
                   glob(d.red));
}

// CalcTorque.F90:36-45 masked wall sums
__global__ __launch_bounds__(RED_THREADS) void k_torque(GridDev g, const double *__restrict__ qt,
                                                       const unsigned char *__restrict__ icyl,
                                                       const unsigned char *__restrict__ ocyl,
                                                       double virotint, double virotext,
                                                       double del, double deln,
                                                       double *__restrict__ part) {
  double s[2] = {0.0, 0.0};
  const int nrm = g.nrm, nr = g.nr;
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < g.P) {
    const double grtlow = ((qt[l + g.P] * g.usrm[2] + qt[l] * g.usrm[1]) * 0.5 - virotint * g.usrc[1]) * del;
    s[0] = grtlow * (double)icyl[l];
    const double grtup = (virotext * g.usrc[nr] - (qt[(long long)(nrm - 2) * g.P + l] * g.usrm[nrm - 1] +
                                                   qt[(long long)(nrm - 1) * g.P + l] * g.usrm[nrm]) * 0.5) * deln;
    s[1] = grtup * (double)ocyl[l];
  }
  const int ops[2] = {0, 0};
  block_reduce<2>(s, ops, part + 2 * blockIdx.x);
}

// CalcVelocityStats.F90:29-48
__global__ __launch_bounds__(RED_THREADS) void k_velstats(GridDev g, const double *__restrict__ qt,
                                                         const double *__restrict__ qr,
                                                         const double *__restrict__ qz,
                                                         double pi, double *__restrict__ part) {
  double v[6] = {-100.0, -100.0, -100.0, 0.0, 0.0, 0.0};
  const Col c = column(g, blockIdx.x);
  if (c.ok) {
    double r_c = qr[c.l];
    for (int kc = 1; kc <= g.nrm; kc++) {
      const long long b = (long long)(kc - 1) * g.P;
      const double a = qt[b + c.l], cz = qz[b + c.l];
      const double r_kp = qr[b + g.P + c.l];
      v[0] = fmax(v[0], fabs(a));
      v[1] = fmax(v[1], fabs(r_c));
      v[2] = fmax(v[2], fabs(cz));
      const double t = a + qt[b + c.lip];
      v[3] += t * t * g.pvol[kc] * pi * 0.25;
      const double u = r_c * g.usrc[kc] + r_kp * g.usrc[kc + 1];
      v[4] += u * u * g.pvol[kc] * pi * 0.25;
      const double w = cz + qz[b + c.ljp];
      v[5] += w * w * g.pvol[kc] * pi * 0.25;
      r_c = r_kp;
    }
  }
  const int ops[6] = {1, 1, 1, 0, 0, 0};
  block_reduce<6>(v, ops, part + 6 * blockIdx.x);
}

// combine: block q reduces value q over the nb partials (fixed order: strided
// per-thread sums, then a fixed tree)
DEVI void combine_body(const int q, const double *__restrict__ part, int nb, int nv,
                       unsigned opmask, double *__restrict__ out) {
  __shared__ double sh[256];
  const bool mx = (opmask >> q) & 1u;
  double x = mx ? -1.0e308 : 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const double y = part[(size_t)b * nv + q];
    x = mx ? fmax(x, y) : x + y;
  }
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double y = sh[threadIdx.x + w];
      sh[threadIdx.x] = mx ? fmax(sh[threadIdx.x], y) : sh[threadIdx.x] + y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = sh[0];
}

__global__ __launch_bounds__(256) void k_combine(const double *__restrict__ part, int nb, int nv,
                                                 unsigned opmask, double *__restrict__ out) {
  combine_body(blockIdx.x, part, nb, nv, opmask, out);
}

// Grouped combine: blockIdx.x = value, blockIdx.y = instance; with dyn, only
// instances logging this step.
__global__ __launch_bounds__(256) void k_combine_g(const StepDesc *__restrict__ tab,
                                                   const BatchDyn *__restrict__ dyn, int nb,
                                                   int nv, unsigned opmask) {
  if (dyn && !dyn[blockIdx.y].log) return;
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  combine_body(blockIdx.x, glob(d.red), nb, nv, opmask, glob(d.red_out));
}

// CalcMaxCFL floor and the fixed-dt guard of main.F90:156,172-173, one
// thread per instance: remember the first step with cfl*dt > CFLMAX.
__global__ void k_cfl_check_g(const StepDesc *__restrict__ tab, const BatchDyn *__restrict__ dyn,
                              BatchRes *__restrict__ res, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const StepDesc &d = tab[q];
  if (inst_stopped(d)) return;
  const double cfl = fmax(1.0e-8, d.red_out[0]);
  res[q].cfl = cfl;
  if (res[q].fail165 == 0 && cfl * d.dt > d.limitcfl) {
    res[q].fail165 = dyn[q].ntime + 1;
    res[q].fail_time = dyn[q].tnew;
  }
}

// Keep the logging pass's values of the instances that logged this step.
__global__ void k_diag_store_g(const StepDesc *__restrict__ tab, const BatchDyn *__restrict__ dyn,
                               BatchRes *__restrict__ res, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || !dyn[q].log || inst_stopped(tab[q])) return;
  const double *o = tab[q].red_out;
  for (int v = 0; v < 10; v++) res[q].diag[v] = o[v];
  res[q].log_ntime = dyn[q].ntime;
  // main.F90:182,203: 266 when both vmax(1) and vmax(2) exceed VMAX, then
  // 169 (assigned later, so it wins) when max|div| > RESID
  int code = 0;
  if (o[2] > tab[q].limitvel && o[3] > tab[q].limitvel) code = 266;
  if (fabs(o[1]) > tab[q].resid) code = 169;
  if (code && res[q].faillog == 0) {
    res[q].faillog = dyn[q].ntime + 1;
    res[q].faillog_code = code;
    if (res[q].fail165 == 0) res[q].fail_time = dyn[q].tnew;
  }
}

static unsigned red_blocks(const tc_ctx *c) { return (unsigned)((c->P + RED_THREADS - 1) / RED_THREADS); }

static void finish_reduction(tc_ctx *c, int nv, unsigned opmask) {
  hipLaunchKernelGGL(k_combine, dim3((unsigned)nv), dim3(256), 0, c->stream, c->red, (int)red_blocks(c),
                     nv, opmask, c->red_out);
  hipMemcpyAsync(c->red_host, c->red_out, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost,
                 c->stream);
}

void launch_max_cfl(tc_ctx *c) {
  timer_begin(c, "max_cfl");
  hipLaunchKernelGGL(k_pointmax<0>, dim3(red_blocks(c)), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->red);
  finish_reduction(c, 1, 1u);
  timer_end(c);
}
void launch_max_div(tc_ctx *c) {
  timer_begin(c, "max_div");
  hipLaunchKernelGGL(k_pointmax<1>, dim3(red_blocks(c)), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->red);
  finish_reduction(c, 1, 1u);
  timer_end(c);
}
// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point tc_check_divergence_noalloc: the
// per-block partials of its divergence reduction live in this fixed device
// array instead of the context's reduction buffer.  Each GPU the library
// runs on holds its own copy, at its own address, until the process ends;
// in host code `scratch` names no device memory, so host code uses only the
// address looked up per device below.  Every block writes its own partial
// before the combine reads it, so nothing in it needs clearing; calls that
// overlap on one GPU race on it.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 65536;  // 512 KiB: planes up to 16 Mi points
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static std::mutex scratch_mu;
static std::map<int, double *> scratch_dev;  // device -> address of scratch there

// The device address of scratch on device dev, looked up once per device
// with that device current.
static int scratch_on(int dev, double **out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  const auto it = scratch_dev.find(dev);
  if (it != scratch_dev.end()) {
    *out = it->second;
    return 0;
  }
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return set_error("hipGetDevice failed");
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return set_error("hipSetDevice failed");
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  scratch_dev[dev] = static_cast<double *>(ptr);
  *out = static_cast<double *>(ptr);
  return 0;
}

int check_divergence_noalloc(tc_ctx *c, double *qmax_dev) {
  const unsigned nb = red_blocks(c);
  if ((size_t)nb > NOALLOC_SCRATCH_DOUBLES)
    return set_error("tc_check_divergence_noalloc: a plane of " + std::to_string(c->P) +
                     " points needs " + std::to_string(nb) + " partials; the fixed scratch holds " +
                     std::to_string(NOALLOC_SCRATCH_DOUBLES) + " (nzm*nthm <= " +
                     std::to_string(NOALLOC_SCRATCH_DOUBLES * RED_THREADS) + ")");
  double *part = nullptr;
  int rc = scratch_on(c->device, &part);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pointmax<1>, dim3(nb), dim3(RED_THREADS), 0, c->stream, c->gd, c->f[TC_QT],
                     c->f[TC_QR], c->f[TC_QZ], part);
  hipLaunchKernelGGL(k_combine, dim3(1u), dim3(256), 0, c->stream, part, (int)nb, 1, 1u, qmax_dev);
  return check_launch();
}

void launch_torque(tc_ctx *c) {
  const double del = -1.0 / (c->g.rc[2] - c->g.rc[1]);
  const double deln = -1.0 / (c->g.rc[c->nr] - c->g.rc[c->nrm]);
  hipLaunchKernelGGL(k_torque, dim3(red_blocks(c)), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->mask, c->mask + c->P, c->p.virotint, c->p.virotext, del,
                     deln, c->red);
  finish_reduction(c, 2, 0u);
}
void launch_velocity_stats(tc_ctx *c) {
  hipLaunchKernelGGL(k_velstats, dim3(red_blocks(c)), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->g.pi, c->red);
  finish_reduction(c, 6, 7u);
}

// One pass over the velocity for the logging step of main.F90:180-186:
// CalcMaxCFL (CalcMaxCFL.F90:36-39), CheckDivergence (CheckDivergence.F90:34-37),
// CalcVelocityStats (CalcVelocityStats.F90:35-44) and CalcTorque
// (CalcTorque.F90:38-43).  v = cfl, |div|, vmax1..3, rms sums 1..3,
// torque sums inner/outer.
DEVI void diagnostics_body(const int blk, const GridDev &g, const double *__restrict__ qt,
                           const double *__restrict__ qr, const double *__restrict__ qz,
                           const unsigned char *__restrict__ icyl,
                           const unsigned char *__restrict__ ocyl, double virotint,
                           double virotext, double del, double deln, double pi,
                           double *__restrict__ part) {
  // no FMA contraction: the per-point values equal those of the fused
  // correction pass (correct_diag_body) and of the CFL pass bit for bit
#pragma clang fp contract(off)
  double v[10] = {0.0, 0.0, -100.0, -100.0, -100.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const Col c = column(g, blk);
  if (c.ok) {
    double r_c = qr[c.l], t_prev = 0.0;
    for (int kc = 1; kc <= g.nrm; kc++) {
      const long long b = (long long)(kc - 1) * g.P;
      const double t_c = qt[b + c.l], t_ip = qt[b + c.lip];
      const double r_kp = qr[b + g.P + c.l];
      const double z_c = qz[b + c.l], z_jp = qz[b + c.ljp];
      const double cfl = (fabs((t_ip + t_c) * 0.5 * g.dxt * g.usrm[kc]) +
                          fabs((r_kp + r_c) * 0.5 * g.udx2rm[kc]) + fabs((z_jp + z_c) * 0.5 * g.dxz));
      const double dv = ((t_ip - t_c) * g.dxt + (r_kp - r_c) * g.udx2m[kc]) * g.usrm[kc] + (z_jp - z_c) * g.dxz;
      v[0] = fmax(v[0], cfl);
      v[1] = fmax(v[1], fabs(dv));
      v[2] = fmax(v[2], fabs(t_c));
      v[3] = fmax(v[3], fabs(r_c));
      v[4] = fmax(v[4], fabs(z_c));
      const double t = t_c + t_ip;
      v[5] += t * t * g.pvol[kc] * pi * 0.25;
      const double u = r_c * g.usrc[kc] + r_kp * g.usrc[kc + 1];
      v[6] += u * u * g.pvol[kc] * pi * 0.25;
      const double w = z_c + z_jp;
      v[7] += w * w * g.pvol[kc] * pi * 0.25;
      if (kc == 2) {
        const double grtlow = ((t_c * g.usrm[2] + t_prev * g.usrm[1]) * 0.5 - virotint * g.usrc[1]) * del;
        v[8] += grtlow * (double)icyl[c.l];
      }
      if (kc == g.nrm) {
        const double grtup = (virotext * g.usrc[g.nr] - (t_prev * g.usrm[g.nrm - 1] + t_c * g.usrm[g.nrm]) * 0.5) * deln;
        v[9] += grtup * (double)ocyl[c.l];
      }
      t_prev = t_c;
      r_c = r_kp;
    }
  }
  const int ops[10] = {1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
  block_reduce<10>(v, ops, part + 10 * blk);
}

__global__ __launch_bounds__(RED_THREADS) void k_diagnostics(
    GridDev g, const double *__restrict__ qt, const double *__restrict__ qr,
    const double *__restrict__ qz, const unsigned char *__restrict__ icyl,
    const unsigned char *__restrict__ ocyl, double virotint, double virotext, double del,
    double deln, double pi, double *__restrict__ part) {
  diagnostics_body(blockIdx.x, g, qt, qr, qz, icyl, ocyl, virotint, virotext, del, deln, pi, part);
}

// Grouped logging pass: blockIdx.y = instance; instances not logging this
// step leave at once.
__global__ __launch_bounds__(RED_THREADS) void k_diagnostics_g(const StepDesc *__restrict__ tab,
                                                              const BatchDyn *__restrict__ dyn) {
  if (!dyn[blockIdx.y].log) return;
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  diagnostics_body(blockIdx.x, glob(d.g), glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]),
                   glob(d.icyl), glob(d.ocyl), d.virotint, d.virotext, d.del, d.deln, d.pi,
                   glob(d.red));
}

void launch_diagnostics(tc_ctx *c) {
  const double del = -1.0 / (c->g.rc[2] - c->g.rc[1]);
  const double deln = -1.0 / (c->g.rc[c->nr] - c->g.rc[c->nrm]);
  timer_begin(c, "diagnostics");
  hipLaunchKernelGGL(k_diagnostics, dim3(red_blocks(c)), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->mask, c->mask + c->P,
                     c->p.virotint, c->p.virotext, del, deln, c->g.pi, c->red);
  finish_reduction(c, 10, 0x1Fu);
  timer_end(c);
}

// ---------------------------------------------------------------------------
// End of a logging step in one pass: the last substep's CorrectVelocity +
// CorrectPressure + qr wall reset (the arithmetic of correct_pair) and the
// logging diagnostics of main.F90:180-186 (those of diagnostics_body) on the
// corrected fields, so the diagnostics read nothing from HBM themselves.
// A thread owns a theta pair (i, i+1) of one z row and marches a segment of
// r-planes: dph(k-1), dph(k) come from the previous plane and qr(k+1) is
// loaded once for the two planes that use it.  The corrected neighbours
// qt(i+2) and qz(j+1) are corrected again where they are needed: qt(i+2)
// comes from the next lane where it holds it, else from the raw field.  So
// that every raw read (neighbours, and qr(k+1) across r segments) sees the
// raw value, the corrected velocity goes to the three spare arrays tout,
// rout, zout (the caller then swaps the field pointers); pr is corrected in
// place (no pressure neighbour is read).  Blocks: 64 pairs x 4 rows, one
// segment of r each; per-block partials of the ten diagnostics in part
// (diagnostics_body order).
// ---------------------------------------------------------------------------
DEVI void correct_diag_body(const int bx, const int by, const int seg, const int nseg,
                            const GridDev &g, const double *__restrict__ qt,
                            const double *__restrict__ qr, const double *__restrict__ qz,
                            double *__restrict__ pr, double *__restrict__ tout,
                            double *__restrict__ rout, double *__restrict__ zout,
                            const double *__restrict__ dph, double dt, double al, double be,
                            const unsigned char *__restrict__ icyl,
                            const unsigned char *__restrict__ ocyl, double virotint,
                            double virotext, double del, double deln, double pi,
                            double *__restrict__ part, int blk) {
#pragma clang fp contract(off)
  double v[10] = {0.0, 0.0, -100.0, -100.0, -100.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int lane = threadIdx.x;
  const int i = 2 * (bx * 64 + lane);
  const int j = by * 4 + threadIdx.y;
  const bool ok = i < g.nthm && j < g.nzm;
  // the lane whose pair is the next one along theta holds qt(i+2)
  const bool next_in_wave = lane < 63 && i + 2 < g.nthm;
  const int k0 = 1 + (int)(((long long)seg * g.nrm) / nseg);
  const int k1 = (int)(((long long)(seg + 1) * g.nrm) / nseg);
  const int ic = ok ? i : 0, jc = ok ? j : 0;
  const int im = wrap_m(ic, g.nthm), ip = wrap_p(ic + 1, g.nthm);
  const int jp = wrap_p(jc, g.nzm), jm = wrap_m(jc, g.nzm);
  const long long l = (long long)jc * g.nthm + ic;
  const long long lim = (long long)jc * g.nthm + im, lip = (long long)jc * g.nthm + ip;
  const long long ljm = (long long)jm * g.nthm + ic, ljp = (long long)jp * g.nthm + ic;
  auto ld2 = [](const double *a, long long x) { return *reinterpret_cast<const double2 *>(a + x); };
  auto st2 = [](double *a, long long x, double u, double w) {
    *reinterpret_cast<double2 *>(a + x) = make_double2(u, w);
  };
  const long long P = g.P;
  // planes k0-1 and k0 of dph, the raw qr(k0), the corrected qt(k0-1)
  double2 dkm = ld2(dph, (long long)((k0 > 1 ? k0 - 1 : 1) - 1) * P + l);
  double2 d = ld2(dph, (long long)(k0 - 1) * P + l);
  double2 rraw = ld2(qr, (long long)(k0 - 1) * P + l);
  double2 tprev = make_double2(0.0, 0.0);
  if (k0 > 1) {
    const double2 tr = ld2(qt, (long long)(k0 - 2) * P + l);
    const double dimm = dph[(long long)(k0 - 2) * P + lim];
    tprev = make_double2(corr_qt(g, k0 - 1, tr.x, dkm.x, dimm, dt, al),
                         corr_qt(g, k0 - 1, tr.y, dkm.y, dkm.x, dt, al));
  }
  for (int kc = k0; kc <= k1; kc++) {
    const long long b = (long long)(kc - 1) * P;
    const int kp = kc < g.nrm ? kc + 1 : g.nrm;
    const double2 dkp = ld2(dph, (long long)(kp - 1) * P + l);
    const double2 djm = ld2(dph, b + ljm), djp = ld2(dph, b + ljp);
    const double dim_ = dph[b + lim], dip = dph[b + lip];
    const double2 t = ld2(qt, b + l), z = ld2(qz, b + l), p = ld2_nt(pr + b + l);
    // qr(k+1): the wall plane nr is not corrected and ends at zero
    const double2 rnext = ld2(qr, b + P + l);
    const double2 tc = make_double2(corr_qt(g, kc, t.x, d.x, dim_, dt, al),
                                    corr_qt(g, kc, t.y, d.y, d.x, dt, al));
    double2 rc = make_double2(corr_qr(g, kc, rraw.x, d.x, dkm.x, dt, al),
                              corr_qr(g, kc, rraw.y, d.y, dkm.y, dt, al));
    const double2 zc = make_double2(corr_qz(g, z.x, d.x, djm.x, dt, al),
                                    corr_qz(g, z.y, d.y, djm.y, dt, al));
    const double2 pc =
        make_double2(corr_pr(g, kc, p.x, d.x, dim_, d.y, djm.x, djp.x, dkm.x, dkp.x, be),
                     corr_pr(g, kc, p.y, d.y, d.x, dip, djm.y, djp.y, dkm.y, dkp.y, be));
    if (ok) {
      st2_nt(tout + b + l, tc.x, tc.y);
      // CorrectVelocity.F90:46-47: qr(1) = qr(nr) = 0
      if (kc == 1)
        st2_nt(rout + b + l, 0.0, 0.0);
      else
        st2_nt(rout + b + l, rc.x, rc.y);
      st2_nt(zout + b + l, zc.x, zc.y);
      st2_nt(pr + b + l, pc.x, pc.y);
      if (kc == g.nrm) {
        // the plane k = nr: qr's wall, and the unused plane of qt and qz
        // carried over as it was
        const long long bn = (long long)g.nrm * P + l;
        st2_nt(rout + bn, 0.0, 0.0);
        const double2 tw = ld2(qt, bn), zw = ld2(qz, bn);
        st2_nt(tout + bn, tw.x, tw.y);
        st2_nt(zout + bn, zw.x, zw.y);
      }
    }
    {
      if (kc == 1) rc = make_double2(0.0, 0.0);
      // corrected neighbours: qr(k+1) (zero at the wall nr), qt(i+2), qz(j+1)
      double2 rkp = make_double2(0.0, 0.0);
      if (kc < g.nrm)
        rkp = make_double2(corr_qr(g, kc + 1, rnext.x, dkp.x, d.x, dt, al),
                           corr_qr(g, kc + 1, rnext.y, dkp.y, d.y, dt, al));
      double tn = __shfl_down(tc.x, 1, 64);
      if (!next_in_wave) tn = corr_qt(g, kc, qt[b + lip], dip, d.y, dt, al);
      const double2 zr = ld2(qz, b + ljp);
      const double2 zn = make_double2(corr_qz(g, zr.x, djp.x, d.x, dt, al),
                                      corr_qz(g, zr.y, djp.y, d.y, dt, al));
      if (ok) {
        const double tt[2] = {tc.x, tc.y}, tip[2] = {tc.y, tn};
        const double rr[2] = {rc.x, rc.y}, rp[2] = {rkp.x, rkp.y};
        const double zz[2] = {zc.x, zc.y}, zp[2] = {zn.x, zn.y};
        const double tpv[2] = {tprev.x, tprev.y};
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const double t_c = tt[e], t_ip = tip[e], r_c = rr[e], r_kp = rp[e], z_c = zz[e],
                       z_jp = zp[e];
          const double cfl = (fabs((t_ip + t_c) * 0.5 * g.dxt * g.usrm[kc]) +
                              fabs((r_kp + r_c) * 0.5 * g.udx2rm[kc]) + fabs((z_jp + z_c) * 0.5 * g.dxz));
          const double dv =
              ((t_ip - t_c) * g.dxt + (r_kp - r_c) * g.udx2m[kc]) * g.usrm[kc] + (z_jp - z_c) * g.dxz;
          v[0] = fmax(v[0], cfl);
          v[1] = fmax(v[1], fabs(dv));
          v[2] = fmax(v[2], fabs(t_c));
          v[3] = fmax(v[3], fabs(r_c));
          v[4] = fmax(v[4], fabs(z_c));
          const double tsum = t_c + t_ip;
          v[5] += tsum * tsum * g.pvol[kc] * pi * 0.25;
          const double u = r_c * g.usrc[kc] + r_kp * g.usrc[kc + 1];
          v[6] += u * u * g.pvol[kc] * pi * 0.25;
          const double w = z_c + z_jp;
          v[7] += w * w * g.pvol[kc] * pi * 0.25;
          if (kc == 2) {
            const double grtlow = ((t_c * g.usrm[2] + tpv[e] * g.usrm[1]) * 0.5 - virotint * g.usrc[1]) * del;
            v[8] += grtlow * (double)icyl[l + e];
          }
          if (kc == g.nrm) {
            const double grtup =
                (virotext * g.usrc[g.nr] - (tpv[e] * g.usrm[g.nrm - 1] + t_c * g.usrm[g.nrm]) * 0.5) * deln;
            v[9] += grtup * (double)ocyl[l + e];
          }
        }
      }
      tprev = tc;
    }
    dkm = d;
    d = dkp;
    rraw = rnext;
  }
  const int ops[10] = {1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
  block_reduce<10>(v, ops, part + 10 * blk);
}

__global__ __launch_bounds__(256) void k_correct_diag(
    GridDev g, const double *__restrict__ qt, const double *__restrict__ qr,
    const double *__restrict__ qz, double *__restrict__ pr, double *__restrict__ tout,
    double *__restrict__ rout, double *__restrict__ zout, const double *__restrict__ dph,
    double dt, double al, double be, const unsigned char *__restrict__ icyl,
    const unsigned char *__restrict__ ocyl, double virotint, double virotext, double del,
    double deln, double pi, double *__restrict__ part) {
  // one-dimensional grid of (theta block, z block, r segment) tiles, theta
  // fastest, handed out XCD by XCD so that z-neighbouring tiles share an L2
  // (the rows of a block run unsynchronised: meeting every few planes, so
  // that neighbours' dph and qz are read while still cached, measured no
  // faster)
  const int nbx = (g.nthm / 2 + 63) / 64, nby = (g.nzm + 3) / 4;
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = blk % nbx, by = (blk / nbx) % nby, seg = blk / (nbx * nby);
  correct_diag_body(bx, by, seg, gridDim.x / (nbx * nby), g, qt, qr, qz, pr, tout, rout, zout, dph,
                    dt, al, be, icyl, ocyl, virotint, virotext, del, deln, pi, part, blk);
}

// Segments of r per column: enough blocks for about six per CU (at
// 512x512x385 three segments of 129 planes ran 1.46-1.51 ms against 1.81 ms
// for two and 1.56-1.61 ms for four to twelve), each segment at least 16
// planes, and the partials within the reduction buffer (TC_CDIAG_SEG
// overrides).
int correct_diag_segments(const tc_ctx *c) {
  const long long cols = (long long)((c->nthm / 2 + 63) / 64) * ((c->nzm + 3) / 4);
  static const int env = [] {
    const char *e = std::getenv("TC_CDIAG_SEG");
    return e ? std::atoi(e) : 0;
  }();
  long long s = env > 0 ? env : (1536 + cols - 1) / cols;
  s = std::min<long long>(s, std::max(1, c->nrm / 16));
  s = std::min<long long>(s, (long long)c->red_cap / (10 * cols));
  return (int)std::max<long long>(s, 1);
}

bool correct_diag_applies(const tc_ctx *c) { return c->nthm % 2 == 0 && c->nrm >= 3; }

// The corrected qt, qr, qz land in the spare arrays rhs, dq, qcap (free at
// the end of a substep), whose pointers then trade places with qt, qr, qz.
void launch_correct_diag(tc_ctx *c) {
  const int nseg = correct_diag_segments(c);
  const dim3 grid((unsigned)(((c->nthm / 2 + 63) / 64) * ((c->nzm + 3) / 4) * nseg));
  const double del = -1.0 / (c->g.rc[2] - c->g.rc[1]);
  const double deln = -1.0 / (c->g.rc[c->nr] - c->g.rc[c->nrm]);
  timer_begin(c, "correct_diag");
  hipLaunchKernelGGL(k_correct_diag, grid, dim3(64, 4), 0, c->stream, c->gd, c->f[TC_QT],
                     c->f[TC_QR], c->f[TC_QZ], c->f[TC_PR], c->f[TC_RHS], c->f[TC_DQ],
                     c->f[TC_QCAP], c->f[TC_DPH], c->dt, c->al, c->beta, c->mask, c->mask + c->P,
                     c->p.virotint, c->p.virotext, del, deln, c->g.pi, c->red);
  std::swap(c->f[TC_QT], c->f[TC_RHS]);
  std::swap(c->f[TC_QR], c->f[TC_DQ]);
  std::swap(c->f[TC_QZ], c->f[TC_QCAP]);
  hipLaunchKernelGGL(k_combine, dim3(10u), dim3(256), 0, c->stream, c->red, (int)grid.x, 10,
                     0x1Fu, c->red_out);
  hipMemcpyAsync(c->red_host, c->red_out, sizeof(double) * 10, hipMemcpyDeviceToHost, c->stream);
  timer_end(c);
}

// Plain streaming copy of one field (nr planes): the achievable-bandwidth
// reference and the calibration point for hardware byte counters.  One
// 16-byte element per thread over a grid that covers the field once.
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src,
                                              double2 *__restrict__ dst, long long n2,
                                              const double *__restrict__ s1, double *__restrict__ d1,
                                              long long n) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n2) dst[q] = src[q];
  if (q == 0 && (n & 1)) d1[n - 1] = s1[n - 1];
}

void launch_copy(tc_ctx *c, int dst, int src) {
  timer_begin(c, "copy");
  const long long n2 = c->N / 2;
  const unsigned blocks = (unsigned)((n2 + 255) / 256 > 0 ? (n2 + 255) / 256 : 1);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, c->stream, (const double2 *)c->f[src],
                     (double2 *)c->f[dst], n2, c->f[src], c->f[dst], c->N);
  timer_end(c);
}

// STREAM triad a = b + s*c over whole fields, two doubles per thread.
__global__ __launch_bounds__(256) void k_triad(double2 *__restrict__ a, const double2 *__restrict__ b,
                                               const double2 *__restrict__ cc, double s, long long n2,
                                               double *__restrict__ a1, const double *__restrict__ b1,
                                               const double *__restrict__ c1, long long n) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n2) {
    const double2 x = b[q], y = cc[q];
    a[q] = make_double2(x.x + s * y.x, x.y + s * y.y);
  }
  if (q == 0 && (n & 1)) a1[n - 1] = b1[n - 1] + s * c1[n - 1];
}

void launch_triad(tc_ctx *c, int a, int b, int cc, double s) {
  timer_begin(c, "triad");
  const long long n2 = c->N / 2;
  const unsigned blocks = (unsigned)((n2 + 255) / 256 > 0 ? (n2 + 255) / 256 : 1);
  hipLaunchKernelGGL(k_triad, dim3(blocks), dim3(256), 0, c->stream, (double2 *)c->f[a],
                     (const double2 *)c->f[b], (const double2 *)c->f[cc], s, n2, c->f[a], c->f[b],
                     c->f[cc], c->N);
  timer_end(c);
}

// ---------------------------------------------------------------------------
// Grouped launches over the instances of one shape group
// ---------------------------------------------------------------------------

template <bool U, int LPT>
static void launch_fwd_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, int ns, int n,
                         hipStream_t st, bool corr) {
  const int ntx = (c->nthm + FT_I - 1) / FT_I, ntz = (c->nzm + 4 * LPT - 1) / (4 * LPT);
  const dim3 grid((unsigned)(ntx * ntz), (unsigned)n);
  if (corr)
    hipLaunchKernelGGL((k_helm_fwd_g<U, LPT, true>), grid, dim3(256), 0, st, tab, dyn, ns, ntx, ntz);
  else
    hipLaunchKernelGGL((k_helm_fwd_g<U, LPT, false>), grid, dim3(256), 0, st, tab, dyn, ns, ntx, ntz);
}

// corr: every instance's descriptor carries a deferred correction
void launch_helmholtz_fused_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, int ns, int n,
                              hipStream_t st, bool corr) {
  if (c->masks_uniform_one) {
    if (c->helm_lpt == 2) launch_fwd_g<true, 2>(c, tab, dyn, ns, n, st, corr);
    else launch_fwd_g<true, 1>(c, tab, dyn, ns, n, st, corr);
  } else {
    if (c->helm_lpt == 2) launch_fwd_g<false, 2>(c, tab, dyn, ns, n, st, corr);
    else launch_fwd_g<false, 1>(c, tab, dyn, ns, n, st, corr);
  }
  hipLaunchKernelGGL(k_helm_bwd_g, dim3(line_blocks(c), (unsigned)n), dim3(256), 0, st, tab);
}

void launch_correct_both_g(tc_ctx *c, const StepDesc *tab, int n, hipStream_t st) {
  if (c->nthm % 2 == 0) {
    const dim3 grid((unsigned)((c->nthm / 2 + 63) / 64), (unsigned)((c->nzm + 3) / 4),
                    (unsigned)(c->nrm * n));
    hipLaunchKernelGGL(k_correct_both2_g, grid, dim3(64, 4), 0, st, tab, c->nrm);
  } else {
    const dim3 grid((unsigned)((c->nthm + 63) / 64), (unsigned)((c->nzm + 3) / 4),
                    (unsigned)(c->nrm * n));
    hipLaunchKernelGGL(k_correct_both_g, grid, dim3(64, 4), 0, st, tab, c->nrm);
  }
}

void launch_cfl_check_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, BatchRes *res, int n,
                        hipStream_t st) {
  const unsigned nb = red_blocks(c);
  hipLaunchKernelGGL(k_pointmax_cfl_g, dim3(nb, (unsigned)n), dim3(RED_THREADS), 0, st, tab);
  hipLaunchKernelGGL(k_combine_g, dim3(1, (unsigned)n), dim3(256), 0, st, tab,
                     (const BatchDyn *)nullptr, (int)nb, 1, 1u);
  hipLaunchKernelGGL(k_cfl_check_g, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, tab, dyn, res,
                     n);
}

void launch_diagnostics_g(tc_ctx *c, const StepDesc *tab, const BatchDyn *dyn, BatchRes *res, int n,
                          hipStream_t st) {
  const unsigned nb = red_blocks(c);
  hipLaunchKernelGGL(k_diagnostics_g, dim3(nb, (unsigned)n), dim3(RED_THREADS), 0, st, tab, dyn);
  hipLaunchKernelGGL(k_combine_g, dim3(10, (unsigned)n), dim3(256), 0, st, tab, dyn, (int)nb, 10,
                     0x1Fu);
  hipLaunchKernelGGL(k_diag_store_g, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, tab, dyn,
                     res, n);
}

}  // namespace tc
