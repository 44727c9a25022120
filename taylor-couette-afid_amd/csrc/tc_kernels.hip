// MI355X (gfx950) kernels for the stencil, Helmholtz and reduction parts of
// the Taylor-Couette RK3 step.  FP64 throughout, wave64.
//
// Device layout: every 3-D field is [r][z][theta] with theta contiguous,
// r-plane k (1-based, k = 1..nr) at offset (k-1)*P, P = nzm*nthm.  The
// reference's periodic halos (InitVariables.F90:26-35, halo_common.f90:67-108)
// become index wrap-around, so no halo update is ever launched.
#include "tc_internal.h"

namespace tc {

#define DEVI __device__ __forceinline__

DEVI long long IX(const GridDev &g, int k, int j, int i) {
  return (long long)(k - 1) * g.P + (long long)j * g.nthm + i;
}
DEVI int wrap_p(int i, int n) { return (i + 1 == n) ? 0 : i + 1; }
DEVI int wrap_m(int i, int n) { return (i == 0) ? n - 1 : i - 1; }

// ---------------------------------------------------------------------------
// Explicit terms, one point per thread.
// ---------------------------------------------------------------------------

// ExplicitTermsQT.F90:107-192 at (kc, j, i)
DEVI double explicit_qt_at(const GridDev &g, const double *__restrict__ qt,
                           const double *__restrict__ qr, const double *__restrict__ qz, int kc,
                           int j, int i) {
  const int ip = wrap_p(i, g.nthm), im = wrap_m(i, g.nthm);
  const int jp = wrap_p(j, g.nzm), jm = wrap_m(j, g.nzm);
  const int kmm = kc > 1 ? kc - 1 : 1, kpp = kc < g.nrm ? kc + 1 : g.nrm, kp = kc + 1;
  const double udxt = g.dxt * 0.25 * g.usrm[kc];
  const double udxtq = g.dxtq / (g.ren * (g.rm[kc] * g.rm[kc]));
  const double udxzq = g.dxzq / g.ren;
  const double udxz = g.dxz * 0.25;
  const double t_c = qt[IX(g, kc, j, i)], t_ip = qt[IX(g, kc, j, ip)], t_im = qt[IX(g, kc, j, im)];
  const double t_jp = qt[IX(g, kc, jp, i)], t_jm = qt[IX(g, kc, jm, i)];
  const double t_kp = qt[IX(g, kpp, j, i)], t_km = qt[IX(g, kmm, j, i)];
  const double r_c = qr[IX(g, kc, j, i)], r_im = qr[IX(g, kc, j, im)];
  const double r_kp = qr[IX(g, kp, j, i)], r_kpim = qr[IX(g, kp, j, im)];
  const double z_c = qz[IX(g, kc, j, i)], z_im = qz[IX(g, kc, j, im)];
  const double z_jp = qz[IX(g, kc, jp, i)], z_jpim = qz[IX(g, kc, jp, im)];
  const double usrc_kp = g.usrc[kp], usrc_k = g.usrc[kc];

  const double h11 = ((t_ip + t_c) * (t_ip + t_c) - (t_c + t_im) * (t_c + t_im)) * udxt;
  const double h12d = ((r_kp + r_kpim) * (t_kp + t_c) - (r_c + r_im) * (t_c + t_km)) * 0.25 * g.udx2rm[kc];
  const double h12n = t_c * g.usrm[kc] * ((r_kp + r_kpim) * usrc_kp + (r_c + r_im) * usrc_k) * 0.25;
  const double h12 = h12d + h12n;
  const double h13 = ((z_jp + z_jpim) * (t_jp + t_c) - (z_c + z_im) * (t_c + t_jm)) * udxz;
  const double d11qt = (t_ip - 2.0 * t_c + t_im) * udxtq;
  const double d22qt = (t_jp - 2.0 * t_c + t_jm) * udxzq;
  const double qre = r_kp * usrc_kp + r_c * usrc_k;
  const double qrw = r_kpim * usrc_kp + r_im * usrc_k;
  const double d11qre = (qre - qrw) * g.udx1vis1[kc] * 0.5;
  const double coriol = -((r_c + r_im) * usrc_k + (r_kp + r_kpim) * usrc_kp) * 0.25 * g.romeg;
  return -(h11 + h12 + h13) + d11qre + d22qt + d11qt + coriol;
}

// ExplicitTermsQR.F90:247-347 at (kc, j, i), kc = 2..nrm
DEVI double explicit_qr_at(const GridDev &g, const double *__restrict__ qt,
                           const double *__restrict__ qr, const double *__restrict__ qz, int kc,
                           int j, int i) {
  const int ip = wrap_p(i, g.nthm), im = wrap_m(i, g.nthm);
  const int jp = wrap_p(j, g.nzm), jm = wrap_m(j, g.nzm);
  const int kmm = kc - 1, kp = kc + 1;
  const double udxt = g.dxt * 0.25 * g.usrc[kc];
  const double udxtq = g.dxtq / (g.ren * (g.rc[kc] * g.rc[kc]));
  const double udxzq = g.dxzq / g.ren;
  const double udxz = g.dxz * 0.25;
  const double t_c = qt[IX(g, kc, j, i)], t_ip = qt[IX(g, kc, j, ip)];
  const double t_m = qt[IX(g, kmm, j, i)], t_mip = qt[IX(g, kmm, j, ip)];
  const double r_c = qr[IX(g, kc, j, i)], r_ip = qr[IX(g, kc, j, ip)], r_im = qr[IX(g, kc, j, im)];
  const double r_jp = qr[IX(g, kc, jp, i)], r_jm = qr[IX(g, kc, jm, i)];
  const double r_kp = qr[IX(g, kp, j, i)], r_km = qr[IX(g, kmm, j, i)];
  const double z_c = qz[IX(g, kc, j, i)], z_jp = qz[IX(g, kc, jp, i)];
  const double z_m = qz[IX(g, kmm, j, i)], z_mjp = qz[IX(g, kmm, jp, i)];

  const double tavg = (t_ip + t_mip + t_c + t_m) * 0.25;
  const double h11n = tavg * tavg;
  const double h21 = ((t_ip + t_mip) * (r_ip + r_c) - (t_c + t_m) * (r_c + r_im)) * udxt;
  const double h22 = ((r_kp * g.usrc[kp] + r_c * g.usrc[kc]) * (r_kp + r_c) -
                      (r_c * g.usrc[kc] + r_km * g.usrc[kmm]) * (r_c + r_km)) *
                     g.udx2c[kc] * 0.25;
  const double h23 = ((z_jp + z_mjp) * (r_jp + r_c) - (z_c + z_m) * (r_c + r_jm)) * udxz;
  const double qte = (t_ip + t_mip);
  const double qtw = (t_c + t_m);
  const double d11qre = -(qte - qtw) * g.udx1vis2[kc];
  const double d11qr = (r_ip - 2.0 * r_c + r_im) * udxtq;
  const double d22qr = (r_jp - 2.0 * r_c + r_jm) * udxzq;
  const double coriol = (t_c + t_ip + t_m + t_mip) * 0.25 * g.rc[kc] * g.romeg;
  return -(h21 + h22 + h23) + h11n + d11qr + d22qr + d11qre + coriol;
}

// ExplicitTermsQZ.F90:395-457 at (kc, j, i)
DEVI double explicit_qz_at(const GridDev &g, const double *__restrict__ qt,
                           const double *__restrict__ qr, const double *__restrict__ qz, int kc,
                           int j, int i) {
  const int ip = wrap_p(i, g.nthm), im = wrap_m(i, g.nthm);
  const int jp = wrap_p(j, g.nzm), jm = wrap_m(j, g.nzm);
  const int kmm = kc > 1 ? kc - 1 : 1, kpp = kc < g.nrm ? kc + 1 : g.nrm, kp = kc + 1;
  const double udxt = g.dxt * 0.25 * g.usrm[kc];
  const double udxtq = g.dxtq / (g.ren * (g.rm[kc] * g.rm[kc]));
  const double udxzq = g.dxzq / g.ren;
  const double udxz = g.dxz * 0.25;
  const double t_c = qt[IX(g, kc, j, i)], t_ip = qt[IX(g, kc, j, ip)];
  const double t_jm = qt[IX(g, kc, jm, i)], t_jmip = qt[IX(g, kc, jm, ip)];
  const double r_c = qr[IX(g, kc, j, i)], r_jm = qr[IX(g, kc, jm, i)];
  const double r_kp = qr[IX(g, kp, j, i)], r_kpjm = qr[IX(g, kp, jm, i)];
  const double z_c = qz[IX(g, kc, j, i)], z_ip = qz[IX(g, kc, j, ip)], z_im = qz[IX(g, kc, j, im)];
  const double z_jp = qz[IX(g, kc, jp, i)], z_jm = qz[IX(g, kc, jm, i)];
  const double z_kp = qz[IX(g, kpp, j, i)], z_km = qz[IX(g, kmm, j, i)];

  const double h31 = ((t_ip + t_jmip) * (z_ip + z_c) - (t_c + t_jm) * (z_c + z_im)) * udxt;
  const double h32 = (((r_kp + r_kpjm) * (z_kp + z_c)) - ((r_c + r_jm) * (z_c + z_km))) * g.udx2rm[kc] * 0.25;
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
  if (WHICH == 0) {
    const int kc = blockIdx.z + 1;
    out[IX(g, kc, j, i)] = explicit_qt_at(g, qt, qr, qz, kc, j, i);
  } else if (WHICH == 1) {
    const int kc = blockIdx.z + 2;
    out[IX(g, kc, j, i)] = explicit_qr_at(g, qt, qr, qz, kc, j, i);
  } else {
    const int kc = blockIdx.z + 1;
    out[IX(g, kc, j, i)] = explicit_qz_at(g, qt, qr, qz, kc, j, i);
  }
}

// All three explicit terms in one pass over the velocity (same arithmetic).
__global__ __launch_bounds__(256) void k_explicit_all(GridDev g, const double *__restrict__ qt,
                                                      const double *__restrict__ qr,
                                                      const double *__restrict__ qz,
                                                      double *__restrict__ dq,
                                                      double *__restrict__ dph,
                                                      double *__restrict__ qcap) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nthm || j >= g.nzm) return;
  const int kc = blockIdx.z + 1;
  const long long o = IX(g, kc, j, i);
  dq[o] = explicit_qt_at(g, qt, qr, qz, kc, j, i);
  if (kc >= 2) dph[o] = explicit_qr_at(g, qt, qr, qz, kc, j, i);
  qcap[o] = explicit_qz_at(g, qt, qr, qz, kc, j, i);
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
void launch_explicit_all(tc_ctx *c) {
  timer_begin(c, "explicit_all");
  hipLaunchKernelGGL(k_explicit_all, point_grid(c, c->nrm), dim3(64, 4), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_DQ], c->f[TC_DPH],
                     c->f[TC_QCAP]);
  timer_end(c);
}

// ---------------------------------------------------------------------------
// Helmholtz: right-hand side assembly fused with the batched tridiagonal
// solve.  One radial line (j, i) per lane; lanes run along theta, so every
// load and store of a plane is coalesced.  The LU factors (dgttrf with partial
// pivoting, row-normalised as in SolveImpEqnUpdate_*) do not depend on (j, i)
// except through the wall masks, so they come from a per-substep table with
// one entry per mask combination.  The forward substitution is streamed while
// the right-hand side is produced; its output y goes to the rhs field and is
// read back by the backward substitution.
// ---------------------------------------------------------------------------

struct LUView {
  const double *dl, *du, *du2, *rd, *sc, *pv;
};
DEVI LUView lu_view(const double *base, int nt) {
  const int n = nt / 6;
  LUView v;
  v.dl = base;
  v.du = base + n;
  v.du2 = base + 2 * n;
  v.rd = base + 3 * n;
  v.sc = base + 4 * n;
  v.pv = base + 5 * n;
  return v;
}

// streamed L-solve step for row r (0-based): f is the right-hand side of row r+1
DEVI void lsolve_step(const LUView &T, int r, double f, double &carry, double &out) {
  if (T.pv[r] == 0.0) {
    out = carry;
    carry = f - T.dl[r] * carry;
  } else {
    out = f;
    carry = carry - T.dl[r] * f;
  }
}

struct ImpArgs {
  double al, ga, ro, dt, alre;
  double vwall_in;  // virotint + vinner_t
  double virotext;
  double bc1;       // 0.5*al*dt*dvinner_t*am1j(1)/ren (SolveImpEqnUpdate_QT.F90:191)
  double udxz;      // dxz*al (ImplicitAndUpdateQZ.F90:25)
};

// ImplicitAndUpdateQT.F90:40-117 + SolveImpEqnUpdate_QT.F90:158-204
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
  const int im = wrap_m(i, g.nthm);
  const int n = g.nrm;
  const int ic = icyl[l], oc = ocyl[l];
  const double fic = (double)ic, foc = (double)oc;
  const LUView T = lu_view(lu + (size_t)(ic + 2 * oc) * nt, nt);
  const long long P = g.P;
  const double dxtal = g.dxt * a.al;

  // k = 1 (inner cylinder rows, ImplicitAndUpdateQT.F90:45-68)
  double qm, q0 = qt[l], qp = qt[l + P];
  double carry, yprev;
  {
    const double app = g.ap1j[1];
    const double acc = (g.ac1j[1] * fic + g.ac1j_fs1 * (1 - fic));
    const double amm = (g.am1j[1] * fic + g.am1j_fs1 * (1 - fic));
    const double d33 = qp * app + q0 * acc + a.vwall_in * amm;
    const double dpx = (pr[l] - pr[l - i + im]) * dxtal * g.usrm[1];
    const double dqv = dq[l];
    const double r = (a.ga * dqv + a.ro * ru1[l] + a.alre * d33 - dpx) * a.dt;
    ru1[l] = dqv;
    carry = (r + a.bc1) * T.sc[0];
  }
  for (int k = 2; k <= n; k++) {
    const long long o = (long long)(k - 1) * P + l;
    qm = q0;
    q0 = qp;
    double d33;
    if (k < n) {
      qp = qt[o + P];
      d33 = qp * g.ap1j[k] + q0 * g.ac1j[k] + qm * g.am1j[k];
    } else {
      const double app = (g.ap1j[k] * foc + g.ap1j_fs2 * (1 - foc));
      const double acc = (g.ac1j[k] * foc + g.ac1j_fs2 * (1 - foc));
      d33 = a.virotext * app + q0 * acc + qm * g.am1j[k];
    }
    const double dpx = (pr[o] - pr[o - i + im]) * dxtal * g.usrm[k];
    const double dqv = dq[o];
    const double r = (a.ga * dqv + a.ro * ru1[o] + a.alre * d33 - dpx) * a.dt;
    ru1[o] = dqv;
    lsolve_step(T, k - 2, r * T.sc[k - 1], carry, yprev);
    y[o - P] = yprev;
  }
  // backward substitution (dgtts2), qt += x
  double x1 = carry * T.rd[n - 1];
  qt[(long long)(n - 1) * P + l] += x1;
  double x0 = (y[(long long)(n - 2) * P + l] - T.du[n - 2] * x1) * T.rd[n - 2];
  qt[(long long)(n - 2) * P + l] += x0;
  for (int r = n - 3; r >= 0; r--) {
    const long long o = (long long)r * P + l;
    const double x = (y[o] - T.du[r] * x0 - T.du2[r] * x1) * T.rd[r];
    qt[o] += x;
    x1 = x0;
    x0 = x;
  }
}

// ImplicitAndUpdateQR.F90:247-279 + SolveImpEqnUpdate_QR.F90:335-349: the
// system has nr rows with identity rows at both walls (zero right-hand side).
__global__ __launch_bounds__(256) void k_implicit_qr(GridDev g, ImpArgs a, double *__restrict__ qr,
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
    const double d33 = qp * g.ap2je[k] + q0 * g.ac2je[k] + qm * g.am2je[k];
    const double udx2 = g.rc[k] * a.al * g.udx2c[k];
    const double dpx = (pr[o] - pr[o - P]) * udx2;
    const double dv = dph[o];
    const double r = (a.ga * dv + a.ro * ru2[o] + a.alre * d33 - dpx) * a.dt;
    ru2[o] = dv;
    lsolve_step(T, k - 2, r * T.sc[k - 1], carry, yprev);
    y[o - P] = yprev;
  }
  lsolve_step(T, n - 2, 0.0, carry, yprev);  // row nr: zero right-hand side
  y[(long long)(n - 2) * P + l] = yprev;
  double x1 = carry * T.rd[n - 1];
  double x0 = (yprev - T.du[n - 2] * x1) * T.rd[n - 2];
  if (n - 2 >= 1) qr[(long long)(n - 2) * P + l] += x0;
  for (int r = n - 3; r >= 1; r--) {
    const long long o = (long long)r * P + l;
    const double x = (y[o] - T.du[r] * x0 - T.du2[r] * x1) * T.rd[r];
    qr[o] += x;
    x1 = x0;
    x0 = x;
  }
  // ImplicitAndUpdateQR.F90:278-279
  qr[l] = 0.0;
  qr[(long long)(n - 1) * P + l] = 0.0;
}

// ImplicitAndUpdateQZ.F90:36-115 + SolveImpEqnUpdate_QZ.F90:155-192.  Row nrm
// is assembled the way the reference leaves it: its history term uses
// ru3(nrm) after the loop has already overwritten it with qcap(nrm).
__global__ __launch_bounds__(256) void k_implicit_qz(GridDev g, ImpArgs a, double *__restrict__ qz,
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
  const int jm = wrap_m(j, g.nzm);
  const long long ljm = (long long)jm * g.nthm + i;
  const int n = g.nrm;
  const int ic = icyl[l], oc = ocyl[l];
  const double fic = (double)ic, foc = (double)oc;
  const LUView T = lu_view(lu + (size_t)(ic + 2 * oc) * nt, nt);
  const long long P = g.P;

  double qm, q0 = qz[l], qp = qz[l + P];
  double carry, yprev;
  {
    const double app = g.ap3j[1];
    const double acc = (g.ac3j[1] * fic + g.ac3j_fs1 * (1 - fic));
    const double amm = (g.am3j[1] * fic + g.am3j_fs1 * (1 - fic));
    const double d33 = qp * app + q0 * acc + q0 * amm;
    const double dpx = (pr[l] - pr[ljm]) * a.udxz;
    const double qv = qcap[l];
    const double r = (a.ga * qv + a.ro * ru3[l] + a.alre * d33 - dpx) * a.dt;
    ru3[l] = qv;
    carry = r * T.sc[0];
  }
  for (int k = 2; k <= n; k++) {
    const long long o = (long long)(k - 1) * P + l;
    qm = q0;
    q0 = qp;
    const double qv = qcap[o];
    double d33, hist;
    if (k < n) {
      qp = qz[o + P];
      d33 = qp * g.ap3j[k] + q0 * g.ac3j[k] + qm * g.am3j[k];
      hist = ru3[o];
    } else {
      const double app = (g.ap3j[k] * foc + g.ap3j_fs2 * (1 - foc));
      const double acc = (g.ac3j[k] * foc + g.ac3j_fs2 * (1 - foc));
      d33 = q0 * app + q0 * acc + qm * g.am3j[k];
      hist = qv;
    }
    const double dpx = (pr[o] - pr[(long long)(k - 1) * P + ljm]) * a.udxz;
    const double r = (a.ga * qv + a.ro * hist + a.alre * d33 - dpx) * a.dt;
    ru3[o] = qv;
    lsolve_step(T, k - 2, r * T.sc[k - 1], carry, yprev);
    y[o - P] = yprev;
  }
  double x1 = carry * T.rd[n - 1];
  qz[(long long)(n - 1) * P + l] += x1;
  double x0 = (y[(long long)(n - 2) * P + l] - T.du[n - 2] * x1) * T.rd[n - 2];
  qz[(long long)(n - 2) * P + l] += x0;
  for (int r = n - 3; r >= 0; r--) {
    const long long o = (long long)r * P + l;
    const double x = (y[o] - T.du[r] * x0 - T.du2[r] * x1) * T.rd[r];
    qz[o] += x;
    x1 = x0;
    x0 = x;
  }
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

void launch_implicit_qt(tc_ctx *c) {
  timer_begin(c, "implicit_qt");
  hipLaunchKernelGGL(k_implicit_qt, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QT], c->f[TC_PR], c->f[TC_DQ], c->f[TC_RU1],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QT), c->lu_nt);
  timer_end(c);
}
void launch_implicit_qr(tc_ctx *c) {
  timer_begin(c, "implicit_qr");
  hipLaunchKernelGGL(k_implicit_qr, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QR], c->f[TC_PR], c->f[TC_DPH], c->f[TC_RU2],
                     c->f[TC_RHS], lu_slot(c, LU_QR), c->lu_nt);
  timer_end(c);
}
void launch_implicit_qz(tc_ctx *c) {
  timer_begin(c, "implicit_qz");
  hipLaunchKernelGGL(k_implicit_qz, dim3(line_blocks(c)), dim3(256), 0, c->stream, c->gd,
                     imp_args(c), c->f[TC_QZ], c->f[TC_PR], c->f[TC_QCAP], c->f[TC_RU3],
                     c->f[TC_RHS], c->mask, c->mask + c->P, lu_slot(c, LU_QZ), c->lu_nt);
  timer_end(c);
}

// ---------------------------------------------------------------------------
// Divergence and projection updates, one point per thread
// ---------------------------------------------------------------------------

DEVI double divergence_at(const GridDev &g, const double *__restrict__ qt,
                          const double *__restrict__ qr, const double *__restrict__ qz, int kc,
                          int j, int i) {
  // CalcLocalDivergence.F90:33-35 (also CheckDivergence.F90:83-85)
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

// CorrectVelocity.F90:249-257 (qr walls zeroed by k_zero_wall_qr)
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

// CorrectPressure.F90:305-315 (amm = apph, be = beta)
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

// qr(1,:,:) = qr(nr,:,:) = 0 (CorrectVelocity.F90:263-264)
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
// Reductions: fixed grid of RED_BLOCKS blocks, per-block partials, then one
// block combines the partials in a fixed order (deterministic sums).
// ---------------------------------------------------------------------------

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

template <int NV>
DEVI void block_reduce(double (&v)[NV], const int *ops /*0 sum, 1 max*/, double *out) {
  __shared__ double sh[NV][RED_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double x = sh[q][0];
      for (int w = 1; w < RED_THREADS / 64; w++) x = ops[q] ? fmax(x, sh[q][w]) : x + sh[q][w];
      out[q] = x;
    }
  }
}

// CalcMaxCFL.F90:30-43 and CheckDivergence.F90:77-89 (WHICH 0 / 1)
template <int WHICH>
__global__ __launch_bounds__(RED_THREADS) void k_pointmax(GridDev g, const double *__restrict__ qt,
                                                         const double *__restrict__ qr,
                                                         const double *__restrict__ qz,
                                                         double *__restrict__ part) {
  double m[1] = {0.0};
  const long long total = (long long)g.nrm * g.P;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int kc = (int)(q / g.P) + 1;
    const long long r = q - (long long)(kc - 1) * g.P;
    const int j = (int)(r / g.nthm), i = (int)(r - (long long)j * g.nthm);
    double val;
    if (WHICH == 0) {
      const int ip = wrap_p(i, g.nthm), jp = wrap_p(j, g.nzm);
      const long long o = q;
      val = (fabs((qt[IX(g, kc, j, ip)] + qt[o]) * 0.5 * g.dxt * g.usrm[kc]) +
             fabs((qr[o + g.P] + qr[o]) * 0.5 * g.udx2rm[kc]) +
             fabs((qz[IX(g, kc, jp, i)] + qz[o]) * 0.5 * g.dxz));
    } else {
      val = fabs(divergence_at(g, qt, qr, qz, kc, j, i));
    }
    m[0] = fmax(m[0], val);
  }
  const int ops[1] = {1};
  block_reduce<1>(m, ops, part + blockIdx.x);
}

// CalcTorque.F90:131-140 masked wall sums and CalcVelocityStats.F90:29-48
__global__ __launch_bounds__(RED_THREADS) void k_torque(GridDev g, const double *__restrict__ qt,
                                                       const unsigned char *__restrict__ icyl,
                                                       const unsigned char *__restrict__ ocyl,
                                                       double virotint, double virotext,
                                                       double del, double deln,
                                                       double *__restrict__ part) {
  double s[2] = {0.0, 0.0};
  const int nrm = g.nrm, nr = g.nr;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < g.P;
       l += (long long)gridDim.x * blockDim.x) {
    const double grtlow = ((qt[l + g.P] * g.usrm[2] + qt[l] * g.usrm[1]) * 0.5 - virotint * g.usrc[1]) * del;
    s[0] += grtlow * (double)icyl[l];
    const double grtup = (virotext * g.usrc[nr] - (qt[(long long)(nrm - 2) * g.P + l] * g.usrm[nrm - 1] +
                                                   qt[(long long)(nrm - 1) * g.P + l] * g.usrm[nrm]) * 0.5) * deln;
    s[1] += grtup * (double)ocyl[l];
  }
  const int ops[2] = {0, 0};
  block_reduce<2>(s, ops, part + 2 * blockIdx.x);
}

__global__ __launch_bounds__(RED_THREADS) void k_velstats(GridDev g, const double *__restrict__ qt,
                                                         const double *__restrict__ qr,
                                                         const double *__restrict__ qz,
                                                         double pi, double *__restrict__ part) {
  double v[6] = {-100.0, -100.0, -100.0, 0.0, 0.0, 0.0};
  const long long total = (long long)g.nrm * g.P;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int kc = (int)(q / g.P) + 1;
    const long long r = q - (long long)(kc - 1) * g.P;
    const int j = (int)(r / g.nthm), i = (int)(r - (long long)j * g.nthm);
    const double a = qt[q], b = qr[q], cz = qz[q];
    v[0] = fmax(v[0], fabs(a));
    v[1] = fmax(v[1], fabs(b));
    v[2] = fmax(v[2], fabs(cz));
    const double t = a + qt[IX(g, kc, j, wrap_p(i, g.nthm))];
    v[3] += t * t * g.pvol[kc] * pi * 0.25;
    const double u = b * g.usrc[kc] + qr[q + g.P] * g.usrc[kc + 1];
    v[4] += u * u * g.pvol[kc] * pi * 0.25;
    const double w = cz + qz[IX(g, kc, wrap_p(j, g.nzm), i)];
    v[5] += w * w * g.pvol[kc] * pi * 0.25;
  }
  const int ops[6] = {1, 1, 1, 0, 0, 0};
  block_reduce<6>(v, ops, part + 6 * blockIdx.x);
}

// combine nv values per block over nb blocks, in block order
__global__ void k_combine(const double *__restrict__ part, int nb, int nv, unsigned opmask,
                          double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= nv) return;
  const bool mx = (opmask >> q) & 1u;
  double x = part[q];
  for (int b = 1; b < nb; b++) {
    const double y = part[(size_t)b * nv + q];
    x = mx ? fmax(x, y) : x + y;
  }
  out[q] = x;
}

static void finish_reduction(tc_ctx *c, int nv, unsigned opmask) {
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(64), 0, c->stream, c->red, RED_BLOCKS, nv, opmask,
                     c->red_out);
  hipMemcpyAsync(c->red_host, c->red_out, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost,
                 c->stream);
}

void launch_max_cfl(tc_ctx *c) {
  timer_begin(c, "max_cfl");
  hipLaunchKernelGGL(k_pointmax<0>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->red);
  finish_reduction(c, 1, 1u);
  timer_end(c);
}
void launch_max_div(tc_ctx *c) {
  timer_begin(c, "max_div");
  hipLaunchKernelGGL(k_pointmax<1>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->red);
  finish_reduction(c, 1, 1u);
  timer_end(c);
}
void launch_torque(tc_ctx *c) {
  const double del = -1.0 / (c->g.rc[2] - c->g.rc[1]);
  const double deln = -1.0 / (c->g.rc[c->nr] - c->g.rc[c->nrm]);
  hipLaunchKernelGGL(k_torque, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->mask, c->mask + c->P, c->p.virotint, c->p.virotext, del,
                     deln, c->red);
  finish_reduction(c, 2, 0u);
}
void launch_velocity_stats(tc_ctx *c) {
  hipLaunchKernelGGL(k_velstats, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->gd,
                     c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->g.pi, c->red);
  finish_reduction(c, 6, 7u);
}

}  // namespace tc
