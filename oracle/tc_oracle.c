/*
 * tc_oracle.c -- CPU restatement of the Taylor-Couette AFiD RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity reference for the MI355X path, never
 * part of it.  Loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.
 *
 * It restates the reference Fortran (rodosti/Taylor-Couette-AFiD) routine by
 * routine, in the reference's own array layout (r fastest, periodic halos in
 * z and theta) and loop order, with the reference's quirks kept:
 *   - ImplicitAndUpdateQZ recomputes kc=nrm with the already-overwritten
 *     history ru3 (ImplicitAndUpdateQZ.F90:67-113);
 *   - CorrectPressure uses apph on both radial neighbours and be=beta
 *     (CorrectPressure.F90:23,40-46);
 *   - SolveImpEqnUpdate_QT scales row 1 with the unblended ac1j(1)
 *     (SolveImpEqnUpdate_QT.F90:65-67);
 *   - CalcStats pairs qt with qz(k,j)+qr(k,j+1) in vtvz_mean
 *     (StatRoutines.F90:59-60);
 *   - CalcDissipation's e12 takes (1/r) d q_r/d theta = d v_r/d theta, with
 *     no 1/r (CalcDissipation.F90:78-80), and rm(nr), never assigned by
 *     CreateGrid, is 0 in udx2mm(nrm).
 * Semantics are those of the non-OpenMP build (udxz = dxz*al and ocylsl as
 * set by SlipBCs in ImplicitAndUpdateQZ).
 *
 * Deviations, both forced:
 *   - fixed-dt mode (IDTV=0) uses dt = DTMAX (the reference never assigns dt
 *     there, main.F90:158-174, and divides by zero);
 *   - the singular (kz,ktheta)=(0,0) Poisson mode is solved on rows 1..nrm-1
 *     with phi(nrm)=0 instead of LAPACK's roundoff-defined last pivot
 *     (SolvePressureCorrection.F90:139-142).
 *
 * Third-party arithmetic restated here: LAPACK dgttrf/dgttrs/zgttrf/zgttrs
 * (reference LAPACK algorithm, partial pivoting) and the unnormalised DFT
 * that FFTW 3.3.4 computes (its own mixed-radix FFT, not FFTW's plans).
 */
#include "tc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[256];
const char *or_last_error(void) { return g_err; }

struct or_state {
  or_params p;
  int nth, nz, nr, nthm, nzm, nrm;
  /* scalars (param.F90:9-80) */
  double pi, time, dt, beta, al, ga, ro;
  double dxt, dxz, dxr, dxtq, dxzq, dxrq;
  double vinner_t, dvinner_t;
  double gam[4], rom[4], alm[4];
  int variabletstep;
  /* 1-D grid vectors, 1-based (param.F90:34-64) */
  double *tc, *tm, *zc, *zm, *rc, *rm, *g2rc, *g2rm, *g2rmm, *g2rcc, *pvol;
  double *usrm, *usrc, *udx1vis1, *udx1vis2, *udx1qm, *udx1qc, *udx2c,
      *udx2m, *udx2cc, *udx2mm, *rudx2c, *udx2rm;
  int *kmc, *kpc, *kmv, *kpv, *kup, *kum;
  double *ap1j, *am1j, *ac1j, *ap2j, *am2j, *ac2j, *ap2je, *am2je, *ac2je,
      *ap3j, *am3j, *ac3j, *apph, *amph, *acph;
  double ap1j_fs[3], ac1j_fs[3], am1j_fs[3], ap3j_fs[3], ac3j_fs[3],
      am3j_fs[3];
  double *ak1, *ap, *ak3, *ao;
  int *icylsl, *ocylsl; /* (nzm, nthm), 1-based via MASK() */
  /* 3-D fields */
  double *f[OR_NFIELDS];
  /* stat_arrays: OR_NSTATS accumulators of (nrm, nzm), Fortran order */
  double *stats;
  int nstatsamples;
  double last_nute;
};

#define HIDX(s, k, j, i) \
  ((size_t)((k) - 1) + (size_t)(s)->nr * ((size_t)(j) + (size_t)((s)->nzm + 2) * (size_t)(i)))
#define NIDX(s, k, j, i) \
  ((size_t)((k) - 1) + (size_t)(s)->nr * ((size_t)((j) - 1) + (size_t)(s)->nzm * (size_t)((i) - 1)))
#define MIDX(s, j, i) ((size_t)((j) - 1) + (size_t)(s)->nzm * (size_t)((i) - 1))

#define QT(k, j, i) s->f[OR_QT][HIDX(s, k, j, i)]
#define QR(k, j, i) s->f[OR_QR][HIDX(s, k, j, i)]
#define QZ(k, j, i) s->f[OR_QZ][HIDX(s, k, j, i)]
#define PR(k, j, i) s->f[OR_PR][HIDX(s, k, j, i)]
#define DPH(k, j, i) s->f[OR_DPH][HIDX(s, k, j, i)]
#define RHS(k, j, i) s->f[OR_RHS][NIDX(s, k, j, i)]
#define DQ(k, j, i) s->f[OR_DQ][NIDX(s, k, j, i)]
#define QCAP(k, j, i) s->f[OR_QCAP][NIDX(s, k, j, i)]
#define RU1(k, j, i) s->f[OR_RU1][NIDX(s, k, j, i)]
#define RU2(k, j, i) s->f[OR_RU2][NIDX(s, k, j, i)]
#define RU3(k, j, i) s->f[OR_RU3][NIDX(s, k, j, i)]
#define ICYL(j, i) s->icylsl[MIDX(s, j, i)]
#define OCYL(j, i) s->ocylsl[MIDX(s, j, i)]

static int is_halo_field(int id) { return id <= OR_DPH; }

static double *dvec(int n) { return (double *)calloc((size_t)n + 2, sizeof(double)); }
static int *ivec(int n) { return (int *)calloc((size_t)n + 2, sizeof(int)); }

/* ------------------------------------------------------------------ */
/* InitTimeMarchScheme.F90:16-48                                      */
/* ------------------------------------------------------------------ */
static void init_time_march_scheme(or_state *s) {
  if (s->p.nsst > 1) {
    s->gam[1] = 8.0 / 15.0;
    s->gam[2] = 5.0 / 12.0;
    s->gam[3] = 3.0 / 4.0;
    s->rom[1] = 0.0;
    s->rom[2] = -17.0 / 60.0;
    s->rom[3] = -5.0 / 12.0;
  } else {
    s->gam[1] = 1.5;
    s->gam[2] = 0.0;
    s->gam[3] = 0.0;
    s->rom[1] = -0.5;
    s->rom[2] = 0.0;
    s->rom[3] = 0.0;
  }
  for (int ns = 1; ns <= s->p.nsst; ns++) s->alm[ns] = (s->gam[ns] + s->rom[ns]);
}

/* ------------------------------------------------------------------ */
/* SlipBCs.F90:31-90 (masks; all ones unless MINSTR*·MULTMIN != 0)      */
/* ------------------------------------------------------------------ */
static void slip_bcs(or_state *s) {
  const int nthm = s->nthm, nzm = s->nzm;
  const double pi = s->pi;
  int acute = 1;
  for (int i = 1; i <= nthm; i++)
    for (int j = 1; j <= nzm; j++) {
      ICYL(j, i) = 1;
      OCYL(j, i) = 1;
    }
  int ntott = s->p.minstrt * s->p.multmin;
  int ntotz = s->p.minstrz * s->p.multmin;
  int ntot = ntott + ntotz;
  double perdist = 2 * pi / s->p.lamb;
  double wint = s->p.virotint / s->p.rint;
  double anglemoved = wint * s->time;
  anglemoved = fmod(anglemoved, perdist);
  int points_moved = (int)floor(anglemoved * s->dxt);
  int *thdispl = (int *)calloc((size_t)nthm + 1, sizeof(int));
  for (int i = 1; i <= nthm; i++) {
    thdispl[i] = (i + points_moved) % nthm;
    if (thdispl[i] == 0) thdispl[i] = nthm;
  }
  if (ntot != 0) {
    for (int i = 1; i <= nthm; i++)
      for (int j = 1; j <= nzm; j++) ICYL(j, i) = 0;
    double m1 = (double)s->p.minstrt / (double)s->p.minstrz * (double)nzm / (double)nthm;
    double y1 = (double)nzm / (double)ntotz;
    double x1 = (double)nthm / (double)ntott;
    int limit1, limit2;
    if (acute == 1 && ntotz != 0) {
      limit1 = -ntott;
      limit2 = ntotz;
    } else {
      limit1 = 0;
      limit2 = ntot;
    }
    for (int ii = limit1; ii <= limit2; ii++) {
      for (int j = 1; j <= nthm; j++) {
        int truej = thdispl[j];
        if (ntotz == 0 && (j - ii * x1 > 0.00001) && (j - ii * x1 - x1 / 2. < 0.0001)) {
          for (int k = 1; k <= nzm; k++) ICYL(k, truej) = 1;
        } else if (ntotz != 0) {
          for (int k = 1; k <= nzm; k++) {
            if ((k - acute * m1 * j - ii * y1 > 0.00001) &&
                (k - acute * m1 * j - ii * y1 - y1 / 2. < 0.00001))
              ICYL(k, truej) = 1;
          }
        }
      }
    }
  }
  free(thdispl);
}

/* ------------------------------------------------------------------ */
/* CreateGrid.F90:24-323                                              */
/* ------------------------------------------------------------------ */
static int create_grid(or_state *s) {
  const int nthm = s->nthm, nzm = s->nzm, nrm = s->nrm, nr = s->nr, nth = s->nth, nz = s->nz;
  const double pi = s->pi;
  const double lamb = s->p.lamb, alx3 = s->p.alx3, rint = s->p.rint, rext = s->p.rext;
  const double str = s->p.str, ren = s->p.ren;
  s->dxt = ((double)nthm * lamb) / (2. * pi);
  s->dxz = (double)nzm / alx3;
  s->dxr = (double)nrm / 1.0;
  s->dxtq = s->dxt * s->dxt;
  s->dxzq = s->dxz * s->dxz;
  s->dxrq = s->dxr * s->dxr;
  const double dxr = s->dxr, dxt = s->dxt, dxtq = s->dxtq, dxrq = s->dxrq;

  for (int kc = 1; kc <= nrm; kc++) {
    s->kmv[kc] = kc - 1;
    s->kpv[kc] = kc + 1;
    if (kc == 1) s->kmv[kc] = kc;
    if (kc == nrm) s->kpv[kc] = kc;
  }
  for (int kc = 1; kc <= nrm; kc++) {
    s->kpc[kc] = s->kpv[kc] - kc;
    s->kmc[kc] = kc - s->kmv[kc];
    s->kup[kc] = 1 - s->kpc[kc];
    s->kum[kc] = 1 - s->kmc[kc];
  }
  for (int i = 1; i <= nth; i++) s->tc[i] = (double)(i - 1) / dxt;
  for (int i = 1; i <= nthm; i++) s->tm[i] = ((double)(i - 1) + 0.5) / dxt;
  for (int j = 1; j <= nz; j++) s->zc[j] = (double)(j - 1) / s->dxz;
  for (int j = 1; j <= nzm; j++) s->zm[j] = ((double)(j - 1) + 0.5) / s->dxz;

  double *rc = s->rc, *rm = s->rm;
  double tstr = tanh(str);
  if (s->p.istr == 0) {
    for (int k = 1; k <= nr; k++) {
      double x3 = (double)(k - 1) / (double)nrm;
      rc[k] = (rext - rint) * x3 + rint;
    }
  } else if (s->p.istr == 4) {
    rc[1] = rint;
    for (int k = 2; k <= nr; k++) {
      double z2dp = (double)(2 * k - nr - 1) / (double)nrm;
      rc[k] = (1 + tanh(str * z2dp) / tstr) * 0.5 * (rext - rint) + rint;
    }
  } else if (s->p.istr == 6) {
    int nclip = (int)str;
    int nrmo = nr + nclip + nclip;
    double *etazm = (double *)calloc((size_t)nrmo + 2, sizeof(double));
    double *etaz = dvec(nr);
    for (int k = 1; k <= nrmo; k++) etazm[k] = +cos(pi * ((double)k - 0.5) / (double)nrmo);
    for (int k = 1; k <= nr; k++) etaz[k] = etazm[k + nclip];
    double delet = etaz[1] - etaz[nr];
    for (int k = 1; k <= nr; k++) etaz[k] = etaz[k] / (0.5 * delet);
    rc[1] = rint;
    for (int k = 2; k <= nrm; k++) rc[k] = (rext - rint) * (1. - etaz[k]) * 0.5 + rint;
    rc[nr] = rext;
    free(etazm);
    free(etaz);
  } else {
    snprintf(g_err, sizeof g_err, "ISTR=%d is not one of 0, 4, 6", s->p.istr);
    return -1;
  }

  double *g2rm = s->g2rm, *g2rc = s->g2rc;
  for (int k = 1; k <= nrm; k++) {
    rm[k] = (rc[k] + rc[k + 1]) * 0.5;
    g2rm[k] = (rc[k + 1] - rc[k]) * dxr;
  }
  g2rc[1] = (rc[2] - rc[1]) * dxr;
  s->g2rmm[1] = (rm[2] - rm[1]) * dxr;
  s->g2rcc[1] = (rc[2] - rc[1]) * dxr;
  for (int k = 2; k <= nrm; k++) {
    g2rc[k] = (rc[k + 1] - rc[k - 1]) * dxr * 0.5;
    s->g2rmm[k] = (rm[k + 1] - rm[k - 1]) * dxr;
    s->g2rcc[k] = (rc[k + 1] - rc[k - 1]) * dxr;
  }
  g2rc[nr] = (rc[nr] - rc[nrm]) * dxr;
  s->g2rmm[nr] = (rm[nr] - rm[nrm]) * dxr;
  s->g2rcc[nr] = (rc[nr] - rc[nrm]) * dxr;

  for (int k = 1; k <= nrm; k++) {
    s->udx1vis1[k] = 2.0 * dxt / (rm[k] * rm[k] * ren);
    s->udx2rm[k] = dxr / (g2rm[k] * rm[k]);
    s->udx2m[k] = dxr / g2rm[k];
    s->udx2mm[k] = dxr / s->g2rmm[k];
    s->udx2c[k] = dxr / g2rc[k];
    s->rudx2c[k] = s->udx2c[k] * rc[k];
    s->usrm[k] = 1.0 / rm[k];
    s->udx1qm[k] = dxtq / (rm[k] * rm[k]);
    s->pvol[k] = (rc[k + 1] * rc[k + 1] - rc[k] * rc[k]);
  }
  s->udx2c[nr] = dxr / g2rc[nr];
  for (int k = 1; k <= nr; k++) {
    s->udx1vis2[k] = dxt / (rc[k] * ren);
    s->usrc[k] = 1.0 / rc[k];
    s->udx1qc[k] = dxtq / (rc[k] * rc[k]);
    s->udx2cc[k] = dxr / s->g2rcc[k];
  }

  /* q1 coefficients, CreateGrid.F90:194-238 */
  for (int kc = 2; kc <= nrm - 1; kc++) {
    int kp = kc + 1;
    double a22 = dxrq / g2rm[kc] / rm[kc];
    double a22p = +a22 * rc[kp] / g2rc[kp];
    double a22m = +a22 * rc[kc] / g2rc[kc];
    s->ap1j[kc] = a22p;
    s->am1j[kc] = a22m;
    s->ac1j[kc] = -(a22p + a22m) - 1. / (rm[kc] * rm[kc]);
  }
  {
    int kc = nrm, kp = kc + 1;
    double a22 = dxrq / (g2rm[kc] * rm[kc]);
    double apjjc = a22 * rc[kp] / g2rc[kp];
    double amjjc = a22 * rc[kc] / g2rc[kc];
    s->ap1j[kc] = 2. * apjjc;
    s->ac1j[kc] = -2. * apjjc - amjjc - 1. / (rm[kc] * rm[kc]);
    s->am1j[kc] = amjjc;
    /* written into index 1, overwritten below; index 2 stays 0 (CreateGrid.F90:218-220) */
    s->ap1j_fs[1] = 0.;
    s->ac1j_fs[1] = -amjjc;
    s->am1j_fs[1] = amjjc;
  }
  {
    int kc = 1, kp = kc + 1;
    double urpp = 1. / (g2rm[kc] * g2rc[kp]);
    double a22 = dxrq / g2rm[kc] / rm[kc];
    double a22p = +a22 * rc[kp] / g2rc[kp];
    double a22m = +a22 * rc[kc] / g2rc[kc];
    s->ap1j[kc] = a22p;
    s->am1j[kc] = 2. * a22m;
    s->ac1j[kc] = -(a22p + 2. * a22m) - 1. / (rm[kc] * rm[kc]);
    s->ap1j_fs[1] = rc[kp] * rc[kp] * rc[kp] / (rm[kc] * rm[kc] * rm[kp]) * dxrq * urpp;
    s->ac1j_fs[1] = -rc[kp] * rc[kp] * rc[kp] / (rm[kc] * rm[kc] * rm[kc]) * dxrq * urpp;
    s->am1j_fs[1] = 0.;
  }
  /* q2 coefficients, CreateGrid.F90:244-259 */
  s->am2j[1] = 0.;
  s->ap2j[1] = 0.;
  s->ac2j[1] = 1.;
  for (int kc = 2; kc <= nrm; kc++) {
    int km = kc - 1;
    double a22 = rc[kc] * dxrq / g2rc[kc];
    double a22p = 1. / (rm[kc] * g2rm[kc]);
    double a22m = 1. / (rm[km] * g2rm[km]);
    s->ap2j[kc] = a22 * a22p;
    s->am2j[kc] = a22 * a22m;
    s->ac2j[kc] = -(a22 * a22p + a22 * a22m);
  }
  s->am2j[nr] = 0.;
  s->ap2j[nr] = 0.;
  s->ac2j[nr] = 1.;
  /* q2 explicit coefficients, CreateGrid.F90:266-276 */
  for (int kc = 2; kc <= nrm; kc++) {
    int km = kc - 1, kp = kc + 1;
    double ugmm = dxrq / g2rc[kc];
    double a22p = rm[kc] / (rc[kp] * g2rm[kc]);
    double a22m = rm[km] / (rc[km] * g2rm[km]);
    double a22 = (rm[kc] / g2rm[kc] + rm[km] / g2rm[km]) / rc[kc];
    s->ap2je[kc] = a22p * ugmm;
    s->am2je[kc] = a22m * ugmm;
    s->ac2je[kc] = -a22 * ugmm - 1. / (rc[kc] * rc[kc]);
  }
  /* q3 coefficients, CreateGrid.F90:281-322 */
  for (int kc = 2; kc <= nrm - 1; kc++) {
    int kp = kc + 1;
    double a22 = dxrq / g2rm[kc] / rm[kc];
    double a22p = +a22 * rc[kp] / g2rc[kp];
    double a22m = +a22 * rc[kc] / g2rc[kc];
    s->ap3j[kc] = a22p;
    s->am3j[kc] = a22m;
    s->ac3j[kc] = -(a22p + a22m);
  }
  {
    int kc = 1, kp = kc + 1;
    double a22 = dxrq / g2rm[kc] / rm[kc];
    double a22p = +a22 * rc[kp] / g2rc[kp];
    double a22m = +a22 * rc[kc] / g2rc[kc];
    s->am3j[kc] = 0.;
    s->ap3j[kc] = a22p;
    s->ac3j[kc] = -(a22p + 2. * a22m);
    s->am3j_fs[1] = 0.;
    s->ap3j_fs[1] = a22p;
    s->ac3j_fs[1] = -(a22p);
  }
  {
    int kc = nrm, kp = kc + 1;
    double a22 = dxrq / g2rm[kc] / rm[kc];
    double a22m = +a22 * rc[kc] / g2rc[kc];
    double a22p = +a22 * rc[kp] / g2rc[kp];
    s->am3j[kc] = a22m;
    s->ap3j[kc] = 0.;
    s->ac3j[kc] = -(a22m + 2. * a22p);
    s->am3j_fs[2] = a22m;
    s->ap3j_fs[2] = 0.;
    s->ac3j_fs[2] = -(a22m);
  }
  slip_bcs(s);
  return 0;
}

/* ------------------------------------------------------------------ */
/* InitPressureSolver.F90:22-65                                       */
/* ------------------------------------------------------------------ */
static void init_pressure_solver(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double pi = s->pi;
  for (int kc = 1; kc <= nrm; kc++) {
    int kp = s->kpv[kc];
    double a22icc = s->rc[kc] * s->kmc[kc] * s->dxrq / s->g2rc[kc];
    double a22icp = s->rc[kp] * s->kpc[kc] * s->dxrq / s->g2rc[kp];
    double ac2 = -(a22icc + a22icp);
    double ugmmm = 1. / (s->rm[kc] * s->g2rm[kc]);
    s->amph[kc] = a22icc * ugmmm;
    s->apph[kc] = a22icp * ugmmm;
    s->acph[kc] = ac2 * ugmmm;
  }
  int nthmh = nthm / 2 + 1, nthmp = nthmh + 1;
  int nzmh = nzm / 2 + 1, nzmp = nzmh + 1;
  for (int j = 1; j <= nzmh; j++) s->ao[j] = (j - 1) * 2. * pi;
  for (int j = nzmp; j <= nzm; j++) s->ao[j] = -((nzm - j + 1) * 2. * pi);
  for (int j = 1; j <= nzm; j++) {
    double t = (double)nzm / s->p.alx3;
    s->ak3[j] = 2. * (1. - cos(s->ao[j] / nzm)) * (t * t);
  }
  for (int i = 1; i <= nthmh; i++) s->ap[i] = (i - 1) * 2.0 * pi;
  for (int i = nthmp; i <= nthm; i++) s->ap[i] = -((nthm - i + 1) * 2.0 * pi);
  for (int i = 1; i <= nthm; i++) s->ak1[i] = 2. * (1. - cos(s->ap[i] / nthm)) * s->dxtq;
}

/* ------------------------------------------------------------------ */
/* creation / field access                                            */
/* ------------------------------------------------------------------ */
or_state *or_create(const or_params *p) {
  if (p->nthm < 2 || p->nzm < 2 || p->nrm < 3) {
    snprintf(g_err, sizeof g_err, "grid too small");
    return NULL;
  }
  if ((p->minstrt * p->multmin) != 0 && p->minstrz == 0) {
    /* SlipBCs.F90:59-61 divides an integer by zero for this input */
    snprintf(g_err, sizeof g_err, "MINSTRT*MULTMIN != 0 with MINSTRZ == 0 is rejected");
    return NULL;
  }
  or_state *s = (or_state *)calloc(1, sizeof(or_state));
  s->p = *p;
  s->nthm = p->nthm;
  s->nzm = p->nzm;
  s->nrm = p->nrm;
  s->nth = p->nthm + 1;
  s->nz = p->nzm + 1;
  s->nr = p->nrm + 1;
  s->pi = 2.0 * asin(1.0);
  s->variabletstep = (p->idtv != 0);
  int nr = s->nr, nth = s->nth, nz = s->nz;
  s->tc = dvec(nth); s->tm = dvec(nth); s->zc = dvec(nz); s->zm = dvec(nz);
  s->rc = dvec(nr); s->rm = dvec(nr); s->g2rc = dvec(nr); s->g2rm = dvec(nr);
  s->g2rmm = dvec(nr); s->g2rcc = dvec(nr); s->pvol = dvec(nr);
  s->usrm = dvec(nr); s->usrc = dvec(nr); s->udx1vis1 = dvec(nr); s->udx1vis2 = dvec(nr);
  s->udx1qm = dvec(nr); s->udx1qc = dvec(nr); s->udx2c = dvec(nr); s->udx2m = dvec(nr);
  s->udx2cc = dvec(nr); s->udx2mm = dvec(nr); s->rudx2c = dvec(nr); s->udx2rm = dvec(nr);
  s->kmc = ivec(nr); s->kpc = ivec(nr); s->kmv = ivec(nr); s->kpv = ivec(nr);
  s->kup = ivec(nr); s->kum = ivec(nr);
  s->ap1j = dvec(nr); s->am1j = dvec(nr); s->ac1j = dvec(nr);
  s->ap2j = dvec(nr); s->am2j = dvec(nr); s->ac2j = dvec(nr);
  s->ap2je = dvec(nr); s->am2je = dvec(nr); s->ac2je = dvec(nr);
  s->ap3j = dvec(nr); s->am3j = dvec(nr); s->ac3j = dvec(nr);
  s->apph = dvec(nr); s->amph = dvec(nr); s->acph = dvec(nr);
  s->ak1 = dvec(nth); s->ap = dvec(nth); s->ak3 = dvec(nz); s->ao = dvec(nz);
  s->icylsl = (int *)calloc((size_t)s->nzm * s->nthm, sizeof(int));
  s->ocylsl = (int *)calloc((size_t)s->nzm * s->nthm, sizeof(int));
  size_t nh = (size_t)nr * (s->nzm + 2) * (s->nthm + 2);
  size_t nn = (size_t)nr * s->nzm * s->nthm;
  for (int id = 0; id < OR_NFIELDS; id++) {
    s->f[id] = (double *)calloc(is_halo_field(id) ? nh : nn, sizeof(double));
    if (!s->f[id]) {
      snprintf(g_err, sizeof g_err, "out of host memory");
      or_destroy(s);
      return NULL;
    }
  }
  s->stats = (double *)calloc((size_t)OR_NSTATS * s->nrm * s->nzm, sizeof(double));
  s->last_nute = NAN;
  init_time_march_scheme(s);
  if (create_grid(s) != 0) {
    or_destroy(s);
    return NULL;
  }
  init_pressure_solver(s);
  return s;
}

void or_destroy(or_state *s) {
  if (!s) return;
  double **v[] = {&s->tc, &s->tm, &s->zc, &s->zm, &s->rc, &s->rm, &s->g2rc, &s->g2rm,
                  &s->g2rmm, &s->g2rcc, &s->pvol, &s->usrm, &s->usrc, &s->udx1vis1,
                  &s->udx1vis2, &s->udx1qm, &s->udx1qc, &s->udx2c, &s->udx2m, &s->udx2cc,
                  &s->udx2mm, &s->rudx2c, &s->udx2rm, &s->ap1j, &s->am1j, &s->ac1j,
                  &s->ap2j, &s->am2j, &s->ac2j, &s->ap2je, &s->am2je, &s->ac2je, &s->ap3j,
                  &s->am3j, &s->ac3j, &s->apph, &s->amph, &s->acph, &s->ak1, &s->ap,
                  &s->ak3, &s->ao};
  for (size_t q = 0; q < sizeof v / sizeof v[0]; q++) free(*v[q]);
  free(s->kmc); free(s->kpc); free(s->kmv); free(s->kpv); free(s->kup); free(s->kum);
  free(s->icylsl); free(s->ocylsl);
  for (int id = 0; id < OR_NFIELDS; id++) free(s->f[id]);
  free(s->stats);
  free(s);
}

int or_get_vec(or_state *s, int id, double *out, int n) {
  const double *v = NULL;
  int len = s->nr;
  switch (id) {
    case OR_V_RC: v = s->rc; break;
    case OR_V_RM: v = s->rm; break;
    case OR_V_G2RC: v = s->g2rc; break;
    case OR_V_G2RM: v = s->g2rm; break;
    case OR_V_USRC: v = s->usrc; break;
    case OR_V_USRM: v = s->usrm; break;
    case OR_V_UDX2RM: v = s->udx2rm; break;
    case OR_V_UDX2M: v = s->udx2m; break;
    case OR_V_UDX2C: v = s->udx2c; break;
    case OR_V_RUDX2C: v = s->rudx2c; break;
    case OR_V_UDX1VIS1: v = s->udx1vis1; break;
    case OR_V_UDX1VIS2: v = s->udx1vis2; break;
    case OR_V_UDX1QM: v = s->udx1qm; break;
    case OR_V_AP1J: v = s->ap1j; break;
    case OR_V_AM1J: v = s->am1j; break;
    case OR_V_AC1J: v = s->ac1j; break;
    case OR_V_AP2J: v = s->ap2j; break;
    case OR_V_AM2J: v = s->am2j; break;
    case OR_V_AC2J: v = s->ac2j; break;
    case OR_V_AP2JE: v = s->ap2je; break;
    case OR_V_AM2JE: v = s->am2je; break;
    case OR_V_AC2JE: v = s->ac2je; break;
    case OR_V_AP3J: v = s->ap3j; break;
    case OR_V_AM3J: v = s->am3j; break;
    case OR_V_AC3J: v = s->ac3j; break;
    case OR_V_APPH: v = s->apph; break;
    case OR_V_AMPH: v = s->amph; break;
    case OR_V_ACPH: v = s->acph; break;
    case OR_V_AK1: v = s->ak1; len = s->nthm; break;
    case OR_V_AK3: v = s->ak3; len = s->nzm; break;
    case OR_V_PVOL: v = s->pvol; break;
    case OR_V_FS: {
      /* ap1j_fs(1:2) ac1j_fs am1j_fs ap3j_fs ac3j_fs am3j_fs, scalars dxt dxz dxr */
      double t[15] = {s->ap1j_fs[1], s->ap1j_fs[2], s->ac1j_fs[1], s->ac1j_fs[2],
                      s->am1j_fs[1], s->am1j_fs[2], s->ap3j_fs[1], s->ap3j_fs[2],
                      s->ac3j_fs[1], s->ac3j_fs[2], s->am3j_fs[1], s->am3j_fs[2],
                      s->dxt, s->dxz, s->dxr};
      int m = n < 15 ? n : 15;
      memcpy(out, t, (size_t)m * sizeof(double));
      return m;
    }
    default: return -1;
  }
  int m = n < len ? n : len;
  for (int q = 0; q < m; q++) out[q] = v[q + 1];
  return m;
}

/* halo_common.f90:67-70 and 105-108: periodic copy on one rank */
void or_update_halo(or_state *s, int id) {
  if (!is_halo_field(id)) return;
  double *a = s->f[id];
  const int nr = s->nr, nzm = s->nzm, nthm = s->nthm;
  for (int i = 0; i <= nthm + 1; i++)
    for (int k = 1; k <= nr; k++) {
      a[HIDX(s, k, nzm + 1, i)] = a[HIDX(s, k, 1, i)];
      a[HIDX(s, k, 0, i)] = a[HIDX(s, k, nzm, i)];
    }
  for (int j = 0; j <= nzm + 1; j++)
    for (int k = 1; k <= nr; k++) {
      a[HIDX(s, k, j, nthm + 1)] = a[HIDX(s, k, j, 1)];
      a[HIDX(s, k, j, 0)] = a[HIDX(s, k, j, nthm)];
    }
}

int or_get_field(or_state *s, int id, double *out) {
  if (id < 0 || id >= OR_NFIELDS) return -1;
  for (int i = 1; i <= s->nthm; i++)
    for (int j = 1; j <= s->nzm; j++)
      for (int k = 1; k <= s->nr; k++)
        out[NIDX(s, k, j, i)] = is_halo_field(id) ? s->f[id][HIDX(s, k, j, i)] : s->f[id][NIDX(s, k, j, i)];
  return 0;
}

int or_set_field(or_state *s, int id, const double *in) {
  if (id < 0 || id >= OR_NFIELDS) return -1;
  for (int i = 1; i <= s->nthm; i++)
    for (int j = 1; j <= s->nzm; j++)
      for (int k = 1; k <= s->nr; k++) {
        if (is_halo_field(id))
          s->f[id][HIDX(s, k, j, i)] = in[NIDX(s, k, j, i)];
        else
          s->f[id][NIDX(s, k, j, i)] = in[NIDX(s, k, j, i)];
      }
  or_update_halo(s, id);
  return 0;
}

void or_get_masks(or_state *s, int *icyl, int *ocyl) {
  size_t n = (size_t)s->nzm * s->nthm;
  memcpy(icyl, s->icylsl, n * sizeof(int));
  memcpy(ocyl, s->ocylsl, n * sizeof(int));
}

void or_set_masks(or_state *s, const int *icyl, const int *ocyl) {
  size_t n = (size_t)s->nzm * s->nthm;
  memcpy(s->icylsl, icyl, n * sizeof(int));
  memcpy(s->ocylsl, ocyl, n * sizeof(int));
}

void or_set_time(or_state *s, double time, double dt) {
  s->time = time;
  s->dt = dt;
}

void or_get_scalars(or_state *s, double *o) {
  o[0] = s->time; o[1] = s->dt; o[2] = s->beta; o[3] = s->al;
  o[4] = s->ga; o[5] = s->ro; o[6] = s->vinner_t; o[7] = s->dvinner_t;
}

/* InnerCylVelocity.F90:15-16 */
static void inner_cyl_velocity(or_state *s) {
  const double pi = s->pi;
  s->vinner_t = s->p.sizepert * sin(2. * pi * s->p.omegapert * s->time);
  s->dvinner_t = 2. * pi * s->p.omegapert * s->p.sizepert * cos(2. * pi * s->p.omegapert * s->time) * s->dt;
}

/* TimeMarcher.F90:20-23 */
void or_begin_step(or_state *s) {
  s->beta = s->dt / s->p.ren * 0.5;
  inner_cyl_velocity(s);
  slip_bcs(s);
}

/* TimeMarcher.F90:27-29 */
void or_set_substep(or_state *s, int ns) {
  s->al = s->alm[ns];
  s->ga = s->gam[ns];
  s->ro = s->rom[ns];
}

/* ------------------------------------------------------------------ */
/* ExplicitTermsQT.F90:22-125                                          */
/* ------------------------------------------------------------------ */
void or_explicit_terms_qt(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double ren = s->p.ren, dxt = s->dxt, dxtq = s->dxtq, romeg = s->p.romeg;
  const double udxzq = s->dxzq / ren;
  const double udxz = s->dxz * 0.25;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kmm = s->kmv[kc], kpp = s->kpv[kc], kp = kc + 1;
        double udxt = dxt * 0.25 * s->usrm[kc];
        double udxtq = dxtq / (ren * (s->rm[kc] * s->rm[kc]));
        double h11 = ((QT(kc, jc, ip) + QT(kc, jc, ic)) * (QT(kc, jc, ip) + QT(kc, jc, ic)) -
                      (QT(kc, jc, ic) + QT(kc, jc, im)) * (QT(kc, jc, ic) + QT(kc, jc, im))) * udxt;
        double h12d = ((QR(kp, jc, ic) + QR(kp, jc, im)) * (QT(kpp, jc, ic) + QT(kc, jc, ic)) -
                       (QR(kc, jc, ic) + QR(kc, jc, im)) * (QT(kc, jc, ic) + QT(kmm, jc, ic))) *
                      0.25 * s->udx2rm[kc];
        double h12n = QT(kc, jc, ic) * s->usrm[kc] *
                      ((QR(kp, jc, ic) + QR(kp, jc, im)) * s->usrc[kp] +
                       (QR(kc, jc, ic) + QR(kc, jc, im)) * s->usrc[kc]) * 0.25;
        double h12 = h12d + h12n;
        double h13 = ((QZ(kc, jp, ic) + QZ(kc, jp, im)) * (QT(kc, jp, ic) + QT(kc, jc, ic)) -
                      (QZ(kc, jc, ic) + QZ(kc, jc, im)) * (QT(kc, jc, ic) + QT(kc, jm, ic))) * udxz;
        double d11qt = (QT(kc, jc, ip) - 2.0 * QT(kc, jc, ic) + QT(kc, jc, im)) * udxtq;
        double d22qt = (QT(kc, jp, ic) - 2.0 * QT(kc, jc, ic) + QT(kc, jm, ic)) * udxzq;
        double qre = QR(kp, jc, ic) * s->usrc[kp] + QR(kc, jc, ic) * s->usrc[kc];
        double qrw = QR(kp, jc, im) * s->usrc[kp] + QR(kc, jc, im) * s->usrc[kc];
        double d11qre = (qre - qrw) * s->udx1vis1[kc] * 0.5;
        double coriol = -((QR(kc, jc, ic) + QR(kc, jc, im)) * s->usrc[kc] +
                          (QR(kp, jc, ic) + QR(kp, jc, im)) * s->usrc[kp]) * 0.25 * romeg;
        DQ(kc, jc, ic) = -(h11 + h12 + h13) + d11qre + d22qt + d11qt + coriol;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* ExplicitTermsQR.F90:25-146  (kc=2..nrm; dph(1) untouched)           */
/* ------------------------------------------------------------------ */
void or_explicit_terms_qr(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double ren = s->p.ren, dxt = s->dxt, dxtq = s->dxtq, romeg = s->p.romeg;
  const double udxzq = s->dxzq / ren;
  const double udxz = s->dxz * 0.25;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 2; kc <= nrm; kc++) {
        double udxt = dxt * 0.25 * s->usrc[kc];
        double udxtq = dxtq / (ren * (s->rc[kc] * s->rc[kc]));
        int kmm = s->kmv[kc], kp = kc + 1;
        double t = (QT(kc, jc, ip) + QT(kmm, jc, ip) + QT(kc, jc, ic) + QT(kmm, jc, ic)) * 0.25;
        double h11n = t * t;
        double h21 = ((QT(kc, jc, ip) + QT(kmm, jc, ip)) * (QR(kc, jc, ip) + QR(kc, jc, ic)) -
                      (QT(kc, jc, ic) + QT(kmm, jc, ic)) * (QR(kc, jc, ic) + QR(kc, jc, im))) * udxt;
        double h22 = ((QR(kp, jc, ic) * s->usrc[kp] + QR(kc, jc, ic) * s->usrc[kc]) *
                          (QR(kp, jc, ic) + QR(kc, jc, ic)) -
                      (QR(kc, jc, ic) * s->usrc[kc] + QR(kmm, jc, ic) * s->usrc[kmm]) *
                          (QR(kc, jc, ic) + QR(kmm, jc, ic))) *
                     s->udx2c[kc] * 0.25;
        double h23 = ((QZ(kc, jp, ic) + QZ(kmm, jp, ic)) * (QR(kc, jp, ic) + QR(kc, jc, ic)) -
                      (QZ(kc, jc, ic) + QZ(kmm, jc, ic)) * (QR(kc, jc, ic) + QR(kc, jm, ic))) * udxz;
        double qte = (QT(kc, jc, ip) + QT(kmm, jc, ip));
        double qtw = (QT(kc, jc, ic) + QT(kmm, jc, ic));
        double d11qre = -(qte - qtw) * s->udx1vis2[kc];
        double d11qr = (QR(kc, jc, ip) - 2.0 * QR(kc, jc, ic) + QR(kc, jc, im)) * udxtq;
        double d22qr = (QR(kc, jp, ic) - 2.0 * QR(kc, jc, ic) + QR(kc, jm, ic)) * udxzq;
        double coriol = (QT(kc, jc, ic) + QT(kc, jc, ip) + QT(kmm, jc, ic) + QT(kmm, jc, ip)) *
                        0.25 * s->rc[kc] * romeg;
        DPH(kc, jc, ic) = -(h21 + h22 + h23) + h11n + d11qr + d22qr + d11qre + coriol;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* ExplicitTermsQZ.F90:23-102                                         */
/* ------------------------------------------------------------------ */
void or_explicit_terms_qz(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double ren = s->p.ren, dxt = s->dxt, dxtq = s->dxtq;
  const double udxz = s->dxz * 0.25;
  const double udxzq = s->dxzq / ren;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kp = kc + 1, kmm = s->kmv[kc], kpp = s->kpv[kc];
        double udxt = dxt * 0.25 * s->usrm[kc];
        double udxtq = dxtq / (ren * (s->rm[kc] * s->rm[kc]));
        double h31 = ((QT(kc, jc, ip) + QT(kc, jm, ip)) * (QZ(kc, jc, ip) + QZ(kc, jc, ic)) -
                      (QT(kc, jc, ic) + QT(kc, jm, ic)) * (QZ(kc, jc, ic) + QZ(kc, jc, im))) * udxt;
        double h32 = (((QR(kp, jc, ic) + QR(kp, jm, ic)) * (QZ(kpp, jc, ic) + QZ(kc, jc, ic))) -
                      ((QR(kc, jc, ic) + QR(kc, jm, ic)) * (QZ(kc, jc, ic) + QZ(kmm, jc, ic)))) *
                     s->udx2rm[kc] * 0.25;
        double h33 = ((QZ(kc, jp, ic) + QZ(kc, jc, ic)) * (QZ(kc, jp, ic) + QZ(kc, jc, ic)) -
                      (QZ(kc, jc, ic) + QZ(kc, jm, ic)) * (QZ(kc, jc, ic) + QZ(kc, jm, ic))) * udxz;
        double dqz1 = (QZ(kc, jc, im) - 2.0 * QZ(kc, jc, ic) + QZ(kc, jc, ip)) * udxtq;
        double dqz2 = (QZ(kc, jm, ic) - 2.0 * QZ(kc, jc, ic) + QZ(kc, jp, ic)) * udxzq;
        QCAP(kc, jc, ic) = -(h31 + h32 + h33) + dqz1 + dqz2;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* LAPACK restatements (reference LAPACK dgttrf/dgtts2, zgttrf/zgtts2) */
/* ------------------------------------------------------------------ */
int or_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv) {
  int info = 0;
  for (int i = 0; i < n; i++) ipiv[i] = i;
  for (int i = 0; i < n - 2; i++) du2[i] = 0.0;
  for (int i = 0; i < n - 2; i++) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      du2[i] = du[i + 1];
      du[i + 1] = -fact * du[i + 1];
      ipiv[i] = i + 1;
    }
  }
  if (n > 1) {
    int i = n - 2;
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      ipiv[i] = i + 1;
    }
  }
  for (int i = 0; i < n; i++)
    if (d[i] == 0.0) {
      info = i + 1;
      break;
    }
  return info;
}

void or_dgttrs(int n, const double *dl, const double *d, const double *du, const double *du2,
               const int *ipiv, double *b) {
  for (int i = 0; i < n - 1; i++) {
    if (ipiv[i] == i) {
      b[i + 1] = b[i + 1] - dl[i] * b[i];
    } else {
      double temp = b[i];
      b[i] = b[i + 1];
      b[i + 1] = temp - dl[i] * b[i];
    }
  }
  b[n - 1] = b[n - 1] / d[n - 1];
  if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
  for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i];
}

/* complex helpers: interleaved (re, im) */
static inline double cabs1(const double *z) { return fabs(z[0]) + fabs(z[1]); }
static inline void cdiv(const double *a, const double *b, double *r) {
  /* Smith's algorithm with range reduction */
  double ar = a[0], ai = a[1], br = b[0], bi = b[1], t, den, rr, ri;
  if (fabs(br) >= fabs(bi)) {
    t = bi / br;
    den = br + bi * t;
    rr = (ar + ai * t) / den;
    ri = (ai - ar * t) / den;
  } else {
    t = br / bi;
    den = br * t + bi;
    rr = (ar * t + ai) / den;
    ri = (ai * t - ar) / den;
  }
  r[0] = rr;
  r[1] = ri;
}
static inline void cmul(const double *a, const double *b, double *r) {
  double rr = a[0] * b[0] - a[1] * b[1];
  double ri = a[0] * b[1] + a[1] * b[0];
  r[0] = rr;
  r[1] = ri;
}

int or_zgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv) {
  int info = 0;
  double fact[2], temp[2], t2[2];
  for (int i = 0; i < n; i++) ipiv[i] = i;
  for (int i = 0; i < n - 2; i++) du2[2 * i] = du2[2 * i + 1] = 0.0;
  for (int i = 0; i < n - 1; i++) {
    double *dli = dl + 2 * i, *di = d + 2 * i, *di1 = d + 2 * (i + 1), *dui = du + 2 * i;
    if (cabs1(di) >= cabs1(dli)) {
      if (cabs1(di) != 0.0) {
        cdiv(dli, di, fact);
        dli[0] = fact[0];
        dli[1] = fact[1];
        cmul(fact, dui, t2);
        di1[0] = di1[0] - t2[0];
        di1[1] = di1[1] - t2[1];
      }
    } else {
      cdiv(di, dli, fact);
      di[0] = dli[0];
      di[1] = dli[1];
      dli[0] = fact[0];
      dli[1] = fact[1];
      temp[0] = dui[0];
      temp[1] = dui[1];
      dui[0] = di1[0];
      dui[1] = di1[1];
      cmul(fact, di1, t2);
      di1[0] = temp[0] - t2[0];
      di1[1] = temp[1] - t2[1];
      if (i < n - 2) {
        double *du1 = du + 2 * (i + 1);
        du2[2 * i] = du1[0];
        du2[2 * i + 1] = du1[1];
        cmul(fact, du1, t2);
        du1[0] = -t2[0];
        du1[1] = -t2[1];
      }
      ipiv[i] = i + 1;
    }
  }
  for (int i = 0; i < n; i++)
    if (cabs1(d + 2 * i) == 0.0) {
      info = i + 1;
      break;
    }
  return info;
}

void or_zgttrs(int n, const double *dl, const double *d, const double *du, const double *du2,
               const int *ipiv, double *b) {
  double t[2], t2[2];
  for (int i = 0; i < n - 1; i++) {
    double *bi = b + 2 * i, *bi1 = b + 2 * (i + 1);
    if (ipiv[i] == i) {
      cmul(dl + 2 * i, bi, t);
      bi1[0] = bi1[0] - t[0];
      bi1[1] = bi1[1] - t[1];
    } else {
      double temp[2] = {bi[0], bi[1]};
      bi[0] = bi1[0];
      bi[1] = bi1[1];
      cmul(dl + 2 * i, bi, t);
      bi1[0] = temp[0] - t[0];
      bi1[1] = temp[1] - t[1];
    }
  }
  cdiv(b + 2 * (n - 1), d + 2 * (n - 1), t);
  b[2 * (n - 1)] = t[0];
  b[2 * (n - 1) + 1] = t[1];
  if (n > 1) {
    cmul(du + 2 * (n - 2), b + 2 * (n - 1), t);
    double num[2] = {b[2 * (n - 2)] - t[0], b[2 * (n - 2) + 1] - t[1]};
    cdiv(num, d + 2 * (n - 2), t2);
    b[2 * (n - 2)] = t2[0];
    b[2 * (n - 2) + 1] = t2[1];
  }
  for (int i = n - 3; i >= 0; i--) {
    double a1[2], a2[2];
    cmul(du + 2 * i, b + 2 * (i + 1), a1);
    cmul(du2 + 2 * i, b + 2 * (i + 2), a2);
    double num[2] = {b[2 * i] - a1[0] - a2[0], b[2 * i + 1] - a1[1] - a2[1]};
    cdiv(num, d + 2 * i, t2);
    b[2 * i] = t2[0];
    b[2 * i + 1] = t2[1];
  }
}

/* ------------------------------------------------------------------ */
/* ImplicitAndUpdateQT.F90:26-121 + SolveImpEqnUpdate_QT.F90:22-78      */
/* ------------------------------------------------------------------ */
static void solve_imp_eqn_update_qt(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double betadx = s->beta * s->al, al = s->al, dt = s->dt, ren = s->p.ren;
#pragma omp parallel
  {
    double *amkl = dvec(nrm), *apkl = dvec(nrm), *ackl = dvec(nrm), *fkl = dvec(nrm);
    double *amkT = dvec(nrm), *ackT = dvec(nrm), *apkT = dvec(nrm), *appk = dvec(nrm);
    int *ipkv = ivec(nrm);
#pragma omp for schedule(static)
    for (int ic = 1; ic <= nthm; ic++) {
      for (int jc = 1; jc <= nzm; jc++) {
        double fic = (double)ICYL(jc, ic), foc = (double)OCYL(jc, ic);
        double actot = (s->ac1j[1] * fic + s->ac1j_fs[1] * (1 - fic));
        double ackl_b = 1.0 / (1.0 - actot * betadx);
        ackl[1] = 1.0;
        apkl[1] = -s->ap1j[1] * betadx * ackl_b;
        amkl[1] = (s->am1j[1] * fic + s->am1j_fs[1] * (1 - fic));
        for (int kc = 2; kc <= nrm - 1; kc++) {
          ackl_b = 1.0 / (1.0 - s->ac1j[kc] * betadx);
          amkl[kc] = -s->am1j[kc] * betadx * ackl_b;
          ackl[kc] = 1.0;
          apkl[kc] = -s->ap1j[kc] * betadx * ackl_b;
        }
        actot = (s->ac1j[nrm] * foc + s->ac1j_fs[2] * (1 - foc));
        ackl_b = 1.0 / (1.0 - actot * betadx);
        amkl[nrm] = -s->am1j[nrm] * betadx * ackl_b;
        ackl[nrm] = 1.0;
        apkl[nrm] = 0.0;
        /* amkT=amkl(2:nrm), apkT=apkl(1:nrm-1), ackT=ackl(1:nrm); 0-based */
        for (int q = 0; q < nrm - 1; q++) {
          amkT[q] = amkl[q + 2];
          apkT[q] = apkl[q + 1];
        }
        for (int q = 0; q < nrm; q++) ackT[q] = ackl[q + 1];
        or_dgttrf(nrm, amkT, ackT, apkT, appk, ipkv);
        ackl_b = 1.0 / (1. - s->ac1j[1] * betadx);
        fkl[0] = (RHS(1, jc, ic) + 0.5 * al * dt * s->dvinner_t * s->am1j[1] / ren) * ackl_b;
        for (int kc = 2; kc <= nrm; kc++) {
          ackl_b = 1.0 / (1. - s->ac1j[kc] * betadx);
          fkl[kc - 1] = RHS(kc, jc, ic) * ackl_b;
        }
        or_dgttrs(nrm, amkT, ackT, apkT, appk, ipkv, fkl);
        for (int kc = 1; kc <= nrm; kc++) QT(kc, jc, ic) = QT(kc, jc, ic) + fkl[kc - 1];
      }
    }
    free(amkl); free(apkl); free(ackl); free(fkl);
    free(amkT); free(ackT); free(apkT); free(appk); free(ipkv);
  }
}

void or_implicit_and_update_qt(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double al = s->al, ga = s->ga, ro = s->ro, dt = s->dt, dxt = s->dxt;
  const double alre = al / s->p.ren;
  const double virotint = s->p.virotint, virotext = s->p.virotext, vinner_t = s->vinner_t;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int imm = ic - 1;
    for (int jc = 1; jc <= nzm; jc++) {
      double fic = (double)ICYL(jc, ic), foc = (double)OCYL(jc, ic);
      int kc = 1, kpp = 2, kmm;
      double app = s->ap1j[kc];
      double acc = (s->ac1j[kc] * fic + s->ac1j_fs[1] * (1 - fic));
      double amm = (s->am1j[kc] * fic + s->am1j_fs[1] * (1 - fic));
      double d33qt = QT(kpp, jc, ic) * app + QT(kc, jc, ic) * acc + (virotint + vinner_t) * amm;
      double dpx11 = (PR(kc, jc, ic) - PR(kc, jc, imm)) * dxt * al * s->usrm[kc];
      RHS(kc, jc, ic) = (ga * DQ(kc, jc, ic) + ro * RU1(kc, jc, ic) + alre * d33qt - dpx11) * dt;
      RU1(kc, jc, ic) = DQ(kc, jc, ic);
      for (kc = 2; kc <= nrm - 1; kc++) {
        kmm = kc - 1;
        kpp = kc + 1;
        amm = s->am1j[kc];
        acc = s->ac1j[kc];
        app = s->ap1j[kc];
        d33qt = QT(kpp, jc, ic) * app + QT(kc, jc, ic) * acc + QT(kmm, jc, ic) * amm;
        dpx11 = (PR(kc, jc, ic) - PR(kc, jc, imm)) * dxt * al * s->usrm[kc];
        RHS(kc, jc, ic) = (ga * DQ(kc, jc, ic) + ro * RU1(kc, jc, ic) + alre * d33qt - dpx11) * dt;
        RU1(kc, jc, ic) = DQ(kc, jc, ic);
      }
      kc = nrm;
      kmm = nrm - 1;
      amm = s->am1j[kc];
      app = (s->ap1j[kc] * foc + s->ap1j_fs[2] * (1 - foc));
      acc = (s->ac1j[kc] * foc + s->ac1j_fs[2] * (1 - foc));
      d33qt = virotext * app + QT(kc, jc, ic) * acc + QT(kmm, jc, ic) * amm;
      dpx11 = (PR(kc, jc, ic) - PR(kc, jc, imm)) * dxt * al * s->usrm[kc];
      RHS(kc, jc, ic) = (ga * DQ(kc, jc, ic) + ro * RU1(kc, jc, ic) + alre * d33qt - dpx11) * dt;
      RU1(kc, jc, ic) = DQ(kc, jc, ic);
    }
  }
  solve_imp_eqn_update_qt(s);
}

/* ------------------------------------------------------------------ */
/* ImplicitAndUpdateQR.F90:28-71 + SolveImpEqnUpdate_QR.F90:25-65       */
/* ------------------------------------------------------------------ */
void or_implicit_and_update_qr(or_state *s) {
  const int nrm = s->nrm, nr = s->nr, nzm = s->nzm, nthm = s->nthm;
  const double al = s->al, ga = s->ga, ro = s->ro, dt = s->dt;
  const double alre = al / s->p.ren;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++)
    for (int jc = 1; jc <= nzm; jc++)
      for (int kc = 2; kc <= nrm; kc++) {
        int kmm = kc - 1, kpp = kc + 1;
        double amm = s->am2je[kc], acc = s->ac2je[kc], app = s->ap2je[kc];
        double udx2 = s->rc[kc] * al * s->udx2c[kc];
        double d33qr = QR(kpp, jc, ic) * app + QR(kc, jc, ic) * acc + QR(kmm, jc, ic) * amm;
        double dpx22 = (PR(kc, jc, ic) - PR(kmm, jc, ic)) * udx2;
        RHS(kc, jc, ic) = (ga * DPH(kc, jc, ic) + ro * RU2(kc, jc, ic) + alre * d33qr - dpx22) * dt;
        RU2(kc, jc, ic) = DPH(kc, jc, ic);
      }
  /* SolveImpEqnUpdate_QR: one factorisation for all lines */
  const double betadx = s->beta * al;
  double *amkl = dvec(nr), *apkl = dvec(nr), *ackl = dvec(nr);
  double *amkT = dvec(nr), *apkT = dvec(nr), *appk = dvec(nr), *ackT = dvec(nr);
  int *ipkv = ivec(nr);
  amkl[1] = 0.;
  apkl[1] = 0.;
  ackl[1] = 1.;
  for (int kc = 2; kc <= nrm; kc++) {
    double ackl_b = 1.0 / (1.0 - s->ac2j[kc] * betadx);
    amkl[kc] = -s->am2j[kc] * betadx * ackl_b;
    ackl[kc] = 1.0;
    apkl[kc] = -s->ap2j[kc] * betadx * ackl_b;
  }
  amkl[nr] = 0.;
  apkl[nr] = 0.;
  ackl[nr] = 1.;
  for (int q = 0; q < nr - 1; q++) {
    amkT[q] = amkl[q + 2];
    apkT[q] = apkl[q + 1];
  }
  for (int q = 0; q < nr; q++) ackT[q] = ackl[q + 1];
  or_dgttrf(nr, amkT, ackT, apkT, appk, ipkv);
#pragma omp parallel
  {
    double *fkl = dvec(nr);
#pragma omp for schedule(static)
    for (int ic = 1; ic <= nthm; ic++)
      for (int jc = 1; jc <= nzm; jc++) {
        fkl[0] = 0.;
        for (int kc = 2; kc <= nrm; kc++) {
          double ackl_b = 1.0 / (1.0 - s->ac2j[kc] * betadx);
          fkl[kc - 1] = RHS(kc, jc, ic) * ackl_b;
        }
        fkl[nr - 1] = 0.;
        or_dgttrs(nr, amkT, ackT, apkT, appk, ipkv, fkl);
        for (int kc = 2; kc <= nrm; kc++) QR(kc, jc, ic) = QR(kc, jc, ic) + fkl[kc - 1];
      }
    free(fkl);
  }
  free(amkl); free(apkl); free(ackl); free(amkT); free(apkT); free(appk); free(ackT); free(ipkv);
  /* ImplicitAndUpdateQR.F90:70-71: whole wall slabs, halos included */
  for (int i = 0; i <= nthm + 1; i++)
    for (int j = 0; j <= nzm + 1; j++) {
      QR(1, j, i) = 0.0;
      QR(nr, j, i) = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* ImplicitAndUpdateQZ.F90:24-119 + SolveImpEqnUpdate_QZ.F90:21-68      */
/* ------------------------------------------------------------------ */
static void solve_imp_eqn_update_qz(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double betadx = s->beta * s->al;
#pragma omp parallel
  {
    double *amkl = dvec(nrm), *apkl = dvec(nrm), *ackl = dvec(nrm), *fkl = dvec(nrm);
    double *amkT = dvec(nrm), *ackT = dvec(nrm), *apkT = dvec(nrm), *appk = dvec(nrm);
    int *ipkv = ivec(nrm);
#pragma omp for schedule(static)
    for (int ic = 1; ic <= nthm; ic++)
      for (int jc = 1; jc <= nzm; jc++) {
        int icy = ICYL(jc, ic), ocy = OCYL(jc, ic);
        double actot = (s->ac3j[1] * icy + s->ac3j_fs[1] * (1 - icy));
        double ackl_b = 1.0 / (1.0 - (actot * betadx));
        ackl[1] = 1.0;
        apkl[1] = -s->ap3j[1] * betadx * ackl_b;
        amkl[1] = 0.0;
        for (int kc = 2; kc <= nrm - 1; kc++) {
          ackl_b = 1.0 / (1.0 - s->ac3j[kc] * betadx);
          amkl[kc] = -s->am3j[kc] * betadx * ackl_b;
          ackl[kc] = 1.0;
          apkl[kc] = -s->ap3j[kc] * betadx * ackl_b;
        }
        actot = (s->ac3j[nrm] * ocy + s->ac3j_fs[2] * (1 - ocy));
        ackl_b = 1.0 / (1.0 - (actot * betadx));
        amkl[nrm] = -s->am3j[nrm] * betadx * ackl_b;
        ackl[nrm] = 1.0;
        apkl[nrm] = 0.0;
        for (int q = 0; q < nrm - 1; q++) {
          amkT[q] = amkl[q + 2];
          apkT[q] = apkl[q + 1];
        }
        for (int q = 0; q < nrm; q++) ackT[q] = ackl[q + 1];
        or_dgttrf(nrm, amkT, ackT, apkT, appk, ipkv);
        for (int kc = 1; kc <= nrm; kc++) {
          ackl_b = 1.0 / (1. - s->ac3j[kc] * betadx);
          fkl[kc - 1] = RHS(kc, jc, ic) * ackl_b;
        }
        or_dgttrs(nrm, amkT, ackT, apkT, appk, ipkv, fkl);
        for (int kc = 1; kc <= nrm; kc++) QZ(kc, jc, ic) = QZ(kc, jc, ic) + fkl[kc - 1];
      }
    free(amkl); free(apkl); free(ackl); free(fkl);
    free(amkT); free(ackT); free(apkT); free(appk); free(ipkv);
  }
}

void or_implicit_and_update_qz(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double al = s->al, ga = s->ga, ro = s->ro, dt = s->dt;
  const double alre = al / s->p.ren;
  const double udxz = s->dxz * al;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++)
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1;
      int icy = ICYL(jc, ic);
      double foc = (double)OCYL(jc, ic);
      int kc = 1, km = s->kmv[kc], kp = s->kpv[kc];
      double app = s->ap3j[kc];
      double acc = (s->ac3j[kc] * icy + s->ac3j_fs[1] * (1 - icy));
      double amm = (s->am3j[kc] * icy + s->am3j_fs[1] * (1 - icy));
      double dqz3 = QZ(kp, jc, ic) * app + QZ(kc, jc, ic) * acc + QZ(km, jc, ic) * amm;
      double dpx33 = (PR(kc, jc, ic) - PR(kc, jm, ic)) * udxz;
      RHS(kc, jc, ic) = (ga * QCAP(kc, jc, ic) + ro * RU3(kc, jc, ic) + alre * dqz3 - dpx33) * dt;
      RU3(kc, jc, ic) = QCAP(kc, jc, ic);
      for (kc = 2; kc <= nrm; kc++) {
        km = s->kmv[kc];
        kp = s->kpv[kc];
        amm = s->am3j[kc];
        acc = s->ac3j[kc];
        app = s->ap3j[kc];
        dqz3 = QZ(kp, jc, ic) * app + QZ(kc, jc, ic) * acc + QZ(km, jc, ic) * amm;
        dpx33 = (PR(kc, jc, ic) - PR(kc, jm, ic)) * udxz;
        RHS(kc, jc, ic) = (ga * QCAP(kc, jc, ic) + ro * RU3(kc, jc, ic) + alre * dqz3 - dpx33) * dt;
        RU3(kc, jc, ic) = QCAP(kc, jc, ic);
      }
      /* outer cylinder: kc=nrm again, amm kept from the loop, ru3(nrm) already
         overwritten (ImplicitAndUpdateQZ.F90:92-113) */
      kc = nrm;
      km = s->kmv[kc];
      kp = s->kpv[kc];
      app = (s->ap3j[kc] * foc + s->ap3j_fs[2] * (1 - foc));
      acc = (s->ac3j[kc] * foc + s->ac3j_fs[2] * (1 - foc));
      dqz3 = QZ(kp, jc, ic) * app + QZ(kc, jc, ic) * acc + QZ(km, jc, ic) * amm;
      dpx33 = (PR(kc, jc, ic) - PR(kc, jm, ic)) * udxz;
      RHS(kc, jc, ic) = (ga * QCAP(kc, jc, ic) + ro * RU3(kc, jc, ic) + alre * dqz3 - dpx33) * dt;
      RU3(kc, jc, ic) = QCAP(kc, jc, ic);
    }
  solve_imp_eqn_update_qz(s);
}

/* ------------------------------------------------------------------ */
/* CalcLocalDivergence.F90:19-40                                      */
/* ------------------------------------------------------------------ */
void or_calc_local_divergence(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double usdt = 1.0 / (s->al * s->dt), dxt = s->dxt, dxz = s->dxz;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kp = kc + 1;
        double dqcap = ((QT(kc, jc, ip) - QT(kc, jc, ic)) * dxt +
                        (QR(kp, jc, ic) - QR(kc, jc, ic)) * s->udx2m[kc]) * s->usrm[kc] +
                       (QZ(kc, jp, ic) - QZ(kc, jc, ic)) * dxz;
        DPH(kc, jc, ic) = dqcap * usdt;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* FFT: unnormalised DFT X[k] = sum_n x[n] exp(sign*2*pi*i*n*k/N)      */
/* (the transform FFTW computes, SolvePressureCorrection.F90:106-113)  */
/* Iterative self-sorting (Stockham) mixed-radix transform over the    */
/* factors of N (4 first, then 2, 3, 5, ...), applied to a batch of B  */
/* interleaved lines: element t of line b is x[2*(t*B + b)].  Twiddles */
/* exp(-2 pi i m/N) come from a long double table.                     */
/* ------------------------------------------------------------------ */
typedef struct {
  int n, nst, rad[64];
  double *tw; /* 2n: cos, -sin of 2 pi m / n */
} fft_plan;

#define OR_MAX_PLANS 32
static fft_plan g_plans[OR_MAX_PLANS];
static int g_nplans = 0;

static const fft_plan *fft_get_plan(int n) {
  const fft_plan *r = NULL;
#pragma omp critical(or_fft_plans)
  {
    for (int q = 0; q < g_nplans; q++)
      if (g_plans[q].n == n) r = &g_plans[q];
    if (!r && g_nplans < OR_MAX_PLANS) {
      fft_plan *pl = &g_plans[g_nplans];
      pl->n = n;
      pl->nst = 0;
      int m = n;
      while (m % 4 == 0) {
        pl->rad[pl->nst++] = 4;
        m /= 4;
      }
      while (m > 1) {
        int p = 2;
        while (m % p) p++;
        pl->rad[pl->nst++] = p;
        m /= p;
      }
      pl->tw = (double *)malloc(2 * (size_t)n * sizeof(double));
      const long double twopi = 6.283185307179586476925286766559005768L;
      for (int q = 0; q < n; q++) {
        long double ang = twopi * (long double)q / (long double)n;
        pl->tw[2 * q] = (double)cosl(ang);
        pl->tw[2 * q + 1] = (double)(-sinl(ang));
      }
      g_nplans++;
      r = pl;
    }
  }
  return r;
}

/* one radix-R pass: sub-transforms of length Ns become length Ns*R */
static void fft_pass(const fft_plan *pl, int R, int Ns, const double *x, double *y, int B, int sign) {
  const int n = pl->n, nq = n / R, step = n / (Ns * R), rstep = n / R;
  const double sg = (sign < 0) ? 1.0 : -1.0; /* table holds exp(-i...) */
  double wr[64], wi[64], vr[64], vi[64], dr[64], di[64];
  double *big = NULL;
  double *Wr = wr, *Wi = wi, *Vr = vr, *Vi = vi, *Dr = dr, *Di = di;
  if (R > 64) {
    big = (double *)malloc(6 * (size_t)R * sizeof(double));
    Wr = big; Wi = big + R; Vr = big + 2 * R; Vi = big + 3 * R; Dr = big + 4 * R; Di = big + 5 * R;
  }
  for (int j = 0; j < nq; j++) {
    const int k = j % Ns;
    const int out0 = (j / Ns) * Ns * R + k;
    for (int r = 0; r < R; r++) {
      const int e = (int)(((long long)k * r * step) % n);
      Wr[r] = pl->tw[2 * e];
      Wi[r] = sg * pl->tw[2 * e + 1];
    }
    for (int b = 0; b < B; b++) {
      for (int r = 0; r < R; r++) {
        const double *v = x + 2 * ((size_t)(j + r * nq) * B + b);
        Vr[r] = Wr[r] * v[0] - Wi[r] * v[1];
        Vi[r] = Wr[r] * v[1] + Wi[r] * v[0];
      }
      if (R == 2) {
        Dr[0] = Vr[0] + Vr[1]; Di[0] = Vi[0] + Vi[1];
        Dr[1] = Vr[0] - Vr[1]; Di[1] = Vi[0] - Vi[1];
      } else if (R == 4) {
        /* w = exp(sign i pi/2) = sign * i */
        const double ar = Vr[0] + Vr[2], ai = Vi[0] + Vi[2];
        const double br = Vr[0] - Vr[2], bi = Vi[0] - Vi[2];
        const double cr = Vr[1] + Vr[3], ci = Vi[1] + Vi[3];
        const double er = Vr[1] - Vr[3], ei = Vi[1] - Vi[3];
        const double s = (sign < 0) ? -1.0 : 1.0; /* w e = s i e */
        const double wer = -s * ei, wei = s * er;
        Dr[0] = ar + cr; Di[0] = ai + ci;
        Dr[1] = br + wer; Di[1] = bi + wei;
        Dr[2] = ar - cr; Di[2] = ai - ci;
        Dr[3] = br - wer; Di[3] = bi - wei;
      } else {
        for (int q = 0; q < R; q++) {
          double sr = 0.0, si = 0.0;
          for (int r = 0; r < R; r++) {
            const int e = (int)(((long long)r * q) % R) * rstep;
            const double cr = pl->tw[2 * e], ci = sg * pl->tw[2 * e + 1];
            sr += cr * Vr[r] - ci * Vi[r];
            si += cr * Vi[r] + ci * Vr[r];
          }
          Dr[q] = sr;
          Di[q] = si;
        }
      }
      for (int r = 0; r < R; r++) {
        double *o = y + 2 * ((size_t)(out0 + r * Ns) * B + b);
        o[0] = Dr[r];
        o[1] = Di[r];
      }
    }
  }
  free(big);
}

/* in-place c2c of B interleaved lines x; ws holds 2 n B doubles */
static void fft_c2c_batch(const fft_plan *pl, double *x, int B, int sign, double *ws) {
  double *a = x, *b = ws;
  int Ns = 1;
  for (int s = 0; s < pl->nst; s++) {
    fft_pass(pl, pl->rad[s], Ns, a, b, B, sign);
    Ns *= pl->rad[s];
    double *t = a;
    a = b;
    b = t;
  }
  if (a != x) memcpy(x, a, 2 * (size_t)pl->n * B * sizeof(double));
}

/* r2c of 2B real lines packed as B complex lines z = x_{2b} + i x_{2b+1}
   (z interleaved as in fft_c2c_batch); out receives the n/2+1 outputs of
   real line c (c = 0..2B-1) at out[2*(t*2B + c)] */
static void fft_r2c_pairs(const fft_plan *pl, double *z, int B, double *out, double *ws) {
  const int n = pl->n, h = n / 2 + 1;
  fft_c2c_batch(pl, z, B, -1, ws);
  for (int t = 0; t < h; t++) {
    const int tm = (n - t) % n;
    for (int b = 0; b < B; b++) {
      const double *p = z + 2 * ((size_t)t * B + b), *q = z + 2 * ((size_t)tm * B + b);
      double *o1 = out + 2 * ((size_t)t * 2 * B + 2 * b), *o2 = o1 + 2;
      /* X1 = (Z(t) + conj Z(n-t))/2, X2 = (Z(t) - conj Z(n-t))/(2i) */
      o1[0] = 0.5 * (p[0] + q[0]);
      o1[1] = 0.5 * (p[1] - q[1]);
      o2[0] = 0.5 * (p[1] + q[1]);
      o2[1] = -0.5 * (p[0] - q[0]);
    }
  }
}

/* c2r of 2B Hermitian half-spectra in (layout of fft_r2c_pairs' out): the
   two real lines 2b, 2b+1 come back as real and imaginary parts of z */
static void fft_c2r_pairs(const fft_plan *pl, const double *in, int B, double *z, double *ws) {
  const int n = pl->n, h = n / 2 + 1;
  for (int t = 0; t < n; t++) {
    const int u = (t < h) ? t : n - t;
    const double cj = (t < h) ? 1.0 : -1.0; /* X(n-t) = conj X(t) */
    for (int b = 0; b < B; b++) {
      const double *x1 = in + 2 * ((size_t)u * 2 * B + 2 * b), *x2 = x1 + 2;
      double *o = z + 2 * ((size_t)t * B + b);
      /* Z = X1 + i X2 */
      o[0] = x1[0] - cj * x2[1];
      o[1] = cj * x1[1] + x2[0];
    }
  }
  fft_c2c_batch(pl, z, B, +1, ws);
}

void or_fft_c2c(int n, double *x, int sign) {
  const fft_plan *pl = fft_get_plan(n);
  double *ws = (double *)malloc(2 * (size_t)n * sizeof(double));
  fft_c2c_batch(pl, x, 1, sign, ws);
  free(ws);
}

void or_fft_r2c(int n, const double *in, double *out) {
  const fft_plan *pl = fft_get_plan(n);
  double *z = (double *)calloc(2 * (size_t)n, sizeof(double));
  double *ws = (double *)malloc(2 * (size_t)n * sizeof(double));
  double *o = (double *)malloc(4 * (size_t)(n / 2 + 1) * sizeof(double));
  for (int t = 0; t < n; t++) z[2 * t] = in[t];
  fft_r2c_pairs(pl, z, 1, o, ws);
  for (int t = 0; t <= n / 2; t++) {
    out[2 * t] = o[4 * t];
    out[2 * t + 1] = o[4 * t + 1];
  }
  free(z); free(ws); free(o);
}

void or_fft_c2r(int n, const double *in, double *out) {
  const fft_plan *pl = fft_get_plan(n);
  const int h = n / 2 + 1;
  double *hin = (double *)calloc(4 * (size_t)h, sizeof(double));
  double *z = (double *)malloc(2 * (size_t)n * sizeof(double));
  double *ws = (double *)malloc(2 * (size_t)n * sizeof(double));
  for (int t = 0; t < h; t++) {
    hin[4 * t] = in[2 * t];
    hin[4 * t + 1] = in[2 * t + 1];
  }
  fft_c2r_pairs(pl, hin, 1, z, ws);
  for (int t = 0; t < n; t++) out[t] = z[2 * t];
  free(hin); free(z); free(ws);
}

/* ------------------------------------------------------------------ */
/* SolvePressureCorrection.F90:100-163                                */
/* layout of the spectrum: (k=1..nrm, kz=1..nzmh, ktheta=1..nthm)     */
/* The transforms run on blocks of OR_KB consecutive radial planes,   */
/* which are contiguous in both the field and the spectrum.           */
/* ------------------------------------------------------------------ */
#define OR_KB 16
void or_solve_pressure_correction(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const int nzmh = nzm / 2 + 1;
  const int nkb = (nrm + OR_KB - 1) / OR_KB;
  size_t ns = (size_t)nrm * nzmh * nthm;
  double *spec = (double *)malloc(2 * ns * sizeof(double));
  const fft_plan *plz = fft_get_plan(nzm), *plt = fft_get_plan(nthm);
#define SP(k, j, i) (spec + 2 * ((size_t)((k) - 1) + (size_t)nrm * ((size_t)((j) - 1) + (size_t)nzmh * (size_t)((i) - 1))))
  /* r2c along z for every (r, theta) (fwd_guruplan_y, :150): planes k0,
     k0+1 of one theta index packed as one complex line */
#pragma omp parallel
  {
    const int B = OR_KB / 2;
    double *z = (double *)malloc(2 * (size_t)nzm * B * sizeof(double));
    double *ws = (double *)malloc(2 * (size_t)nzm * B * sizeof(double));
    double *o = (double *)malloc(4 * (size_t)nzmh * B * sizeof(double));
#pragma omp for schedule(static)
    for (int blk = 0; blk < nthm * nkb; blk++) {
      const int i = blk / nkb + 1, k0 = (blk % nkb) * OR_KB + 1;
      const int nk = (nrm - k0 + 1 < OR_KB) ? nrm - k0 + 1 : OR_KB;
      for (int j = 1; j <= nzm; j++)
        for (int b = 0; b < B; b++) {
          const int ka = k0 + 2 * b, kb = ka + 1;
          z[2 * ((size_t)(j - 1) * B + b)] = (ka - k0 < nk) ? DPH(ka, j, i) : 0.0;
          z[2 * ((size_t)(j - 1) * B + b) + 1] = (kb - k0 < nk) ? DPH(kb, j, i) : 0.0;
        }
      fft_r2c_pairs(plz, z, B, o, ws);
      for (int j = 1; j <= nzmh; j++)
        for (int c = 0; c < nk; c++) {
          SP(k0 + c, j, i)[0] = o[2 * ((size_t)(j - 1) * 2 * B + c)];
          SP(k0 + c, j, i)[1] = o[2 * ((size_t)(j - 1) * 2 * B + c) + 1];
        }
    }
    free(z); free(ws); free(o);
  }
  /* c2c forward along theta for every (r, kz) (fwd_guruplan_z, :154), then
     divide by nthm*nzm (:157); c2c backward (:200) after the solve */
  const double norm = (double)(nthm * nzm);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      /* complex tridiagonal per (kz=j, ktheta=i), :171-192 */
#pragma omp parallel
      {
        double *amphT = (double *)malloc(2 * (size_t)nrm * sizeof(double));
        double *apphT = (double *)malloc(2 * (size_t)nrm * sizeof(double));
        double *acphT = (double *)malloc(2 * (size_t)nrm * sizeof(double));
        double *appph = (double *)malloc(2 * (size_t)nrm * sizeof(double));
        double *drhs = (double *)malloc(2 * (size_t)nrm * sizeof(double));
        int *phpiv = (int *)malloc((size_t)nrm * sizeof(int));
#pragma omp for schedule(static)
        for (int i = 1; i <= nthm; i++)
          for (int j = 1; j <= nzmh; j++) {
            for (int k = 1; k <= nrm; k++) {
              drhs[2 * (k - 1)] = SP(k, j, i)[0];
              drhs[2 * (k - 1) + 1] = SP(k, j, i)[1];
              acphT[2 * (k - 1)] = s->acph[k] - s->ak3[j] - s->ak1[i] * (s->usrm[k] * s->usrm[k]);
              acphT[2 * (k - 1) + 1] = 0.0;
            }
            for (int k = 1; k <= nrm - 1; k++) {
              amphT[2 * (k - 1)] = s->amph[k + 1];
              amphT[2 * (k - 1) + 1] = 0.0;
              apphT[2 * (k - 1)] = s->apph[k];
              apphT[2 * (k - 1) + 1] = 0.0;
            }
            int n = nrm;
            if (i == 1 && j == 1) {
              /* singular Neumann mode: phi(nrm) = 0 gauge, rows 1..nrm-1 */
              n = nrm - 1;
              drhs[2 * (nrm - 1)] = 0.0;
              drhs[2 * (nrm - 1) + 1] = 0.0;
            }
            or_zgttrf(n, amphT, acphT, apphT, appph, phpiv);
            or_zgttrs(n, amphT, acphT, apphT, appph, phpiv, drhs);
            for (int k = 1; k <= nrm; k++) {
              SP(k, j, i)[0] = drhs[2 * (k - 1)];
              SP(k, j, i)[1] = drhs[2 * (k - 1) + 1];
            }
          }
        free(amphT); free(apphT); free(acphT); free(appph); free(drhs); free(phpiv);
      }
    }
    const int sign = pass == 0 ? -1 : +1;
#pragma omp parallel
    {
      double *cl = (double *)malloc(2 * (size_t)nthm * OR_KB * sizeof(double));
      double *ws = (double *)malloc(2 * (size_t)nthm * OR_KB * sizeof(double));
#pragma omp for schedule(static)
      for (int blk = 0; blk < nzmh * nkb; blk++) {
        const int j = blk / nkb + 1, k0 = (blk % nkb) * OR_KB + 1;
        const int nk = (nrm - k0 + 1 < OR_KB) ? nrm - k0 + 1 : OR_KB;
        for (int i = 1; i <= nthm; i++)
          for (int c = 0; c < OR_KB; c++) {
            double *d = cl + 2 * ((size_t)(i - 1) * OR_KB + c);
            if (c < nk) {
              d[0] = SP(k0 + c, j, i)[0];
              d[1] = SP(k0 + c, j, i)[1];
            } else {
              d[0] = d[1] = 0.0;
            }
          }
        fft_c2c_batch(plt, cl, OR_KB, sign, ws);
        for (int i = 1; i <= nthm; i++)
          for (int c = 0; c < nk; c++) {
            const double *d = cl + 2 * ((size_t)(i - 1) * OR_KB + c);
            if (pass == 0) {
              SP(k0 + c, j, i)[0] = d[0] / norm;
              SP(k0 + c, j, i)[1] = d[1] / norm;
            } else {
              SP(k0 + c, j, i)[0] = d[0];
              SP(k0 + c, j, i)[1] = d[1];
            }
          }
      }
      free(cl); free(ws);
    }
  }
  /* c2r along z into dph (:204-207) */
#pragma omp parallel
  {
    const int B = OR_KB / 2;
    double *hin = (double *)malloc(4 * (size_t)nzmh * B * sizeof(double));
    double *z = (double *)malloc(2 * (size_t)nzm * B * sizeof(double));
    double *ws = (double *)malloc(2 * (size_t)nzm * B * sizeof(double));
#pragma omp for schedule(static)
    for (int blk = 0; blk < nthm * nkb; blk++) {
      const int i = blk / nkb + 1, k0 = (blk % nkb) * OR_KB + 1;
      const int nk = (nrm - k0 + 1 < OR_KB) ? nrm - k0 + 1 : OR_KB;
      for (int j = 1; j <= nzmh; j++)
        for (int c = 0; c < 2 * B; c++) {
          double *d = hin + 2 * ((size_t)(j - 1) * 2 * B + c);
          if (c < nk) {
            d[0] = SP(k0 + c, j, i)[0];
            d[1] = SP(k0 + c, j, i)[1];
          } else {
            d[0] = d[1] = 0.0;
          }
        }
      fft_c2r_pairs(plz, hin, B, z, ws);
      for (int j = 1; j <= nzm; j++)
        for (int c = 0; c < nk; c++)
          DPH(k0 + c, j, i) = z[2 * ((size_t)(j - 1) * B + c / 2) + (c & 1)];
    }
    free(hin); free(z); free(ws);
  }
#undef SP
  free(spec);
}

/* ------------------------------------------------------------------ */
/* CorrectVelocity.F90:28-47                                          */
/* ------------------------------------------------------------------ */
void or_correct_velocity(or_state *s) {
  const int nrm = s->nrm, nr = s->nr, nzm = s->nzm, nthm = s->nthm;
  const double dt = s->dt, al = s->al, dxt = s->dxt, dxz = s->dxz;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int im = ic - 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int km = s->kmv[kc];
        double locdph = DPH(kc, jc, ic);
        QT(kc, jc, ic) = QT(kc, jc, ic) - (locdph - DPH(kc, jc, im)) * dt * al * dxt * s->usrm[kc];
        QR(kc, jc, ic) = QR(kc, jc, ic) - (locdph - DPH(km, jc, ic)) * dt * al * s->rudx2c[kc];
        QZ(kc, jc, ic) = QZ(kc, jc, ic) - (locdph - DPH(kc, jm, ic)) * al * dt * dxz;
      }
    }
  }
  for (int i = 0; i <= nthm + 1; i++)
    for (int j = 0; j <= nzm + 1; j++) {
      QR(1, j, i) = 0.0;
      QR(nr, j, i) = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* CorrectPressure.F90:23-46   (amm = apph, be = beta)                  */
/* ------------------------------------------------------------------ */
void or_correct_pressure(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double be = s->beta, dxzq = s->dxzq;
#pragma omp parallel for schedule(static)
  for (int ic = 1; ic <= nthm; ic++) {
    int im = ic - 1, ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jm = jc - 1, jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kp = s->kpv[kc], km = s->kmv[kc];
        double amm = s->apph[kc], acc = s->acph[kc], app = s->apph[kc];
        PR(kc, jc, ic) = PR(kc, jc, ic) + DPH(kc, jc, ic) -
                         be * ((DPH(kc, jc, ip) - 2.0 * DPH(kc, jc, ic) + DPH(kc, jc, im)) * s->udx1qm[kc] +
                               (DPH(kc, jp, ic) - 2.0 * DPH(kc, jc, ic) + DPH(kc, jm, ic)) * dxzq +
                               (DPH(kp, jc, ic) * app + DPH(kc, jc, ic) * acc + DPH(km, jc, ic) * amm));
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* TimeMarcher.F90:25-57 (substep loop)                                */
/* ------------------------------------------------------------------ */
void or_time_marcher(or_state *s) {
  or_begin_step(s);
  for (int ns = 1; ns <= s->p.nsst; ns++) {
    or_set_substep(s, ns);
    or_explicit_terms_qt(s);
    or_explicit_terms_qr(s);
    or_explicit_terms_qz(s);
    or_implicit_and_update_qt(s);
    or_implicit_and_update_qr(s);
    or_implicit_and_update_qz(s);
    or_update_halo(s, OR_QT);
    or_update_halo(s, OR_QR);
    or_update_halo(s, OR_QZ);
    or_calc_local_divergence(s);
    or_solve_pressure_correction(s);
    or_update_halo(s, OR_DPH);
    or_correct_velocity(s);
    or_correct_pressure(s);
    or_update_halo(s, OR_QT);
    or_update_halo(s, OR_QR);
    or_update_halo(s, OR_QZ);
    or_update_halo(s, OR_PR);
  }
}

void or_project(or_state *s) {
  double al = s->al, dt = s->dt;
  s->al = 1.0;
  s->dt = 1.0;
  or_calc_local_divergence(s);
  or_solve_pressure_correction(s);
  or_update_halo(s, OR_DPH);
  or_correct_velocity(s);
  or_update_halo(s, OR_QT);
  or_update_halo(s, OR_QR);
  or_update_halo(s, OR_QZ);
  s->al = al;
  s->dt = dt;
}

/* ------------------------------------------------------------------ */
/* CalcMaxCFL.F90:22-46                                               */
/* ------------------------------------------------------------------ */
double or_calc_max_cfl(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double dxt = s->dxt, dxz = s->dxz;
  double cflm = 0.00000001;
#pragma omp parallel for schedule(static) reduction(max : cflm)
  for (int i = 1; i <= nthm; i++) {
    int ip = i + 1;
    for (int j = 1; j <= nzm; j++) {
      int jp = j + 1;
      for (int k = 1; k <= nrm; k++) {
        int kp = k + 1;
        double qcf = (fabs((QT(k, j, ip) + QT(k, j, i)) * 0.5 * dxt * s->usrm[k]) +
                      fabs((QR(kp, j, i) + QR(k, j, i)) * 0.5 * s->udx2rm[k]) +
                      fabs((QZ(k, jp, i) + QZ(k, j, i)) * 0.5 * dxz));
        cflm = cflm > qcf ? cflm : qcf;
      }
    }
  }
  return cflm;
}

/* ------------------------------------------------------------------ */
/* CheckDivergence.F90:21-43                                          */
/* ------------------------------------------------------------------ */
double or_check_divergence(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double dxt = s->dxt, dxz = s->dxz;
  double qmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : qmax)
  for (int ic = 1; ic <= nthm; ic++) {
    int ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kp = kc + 1;
        double dqcap = ((QT(kc, jc, ip) - QT(kc, jc, ic)) * dxt +
                        (QR(kp, jc, ic) - QR(kc, jc, ic)) * s->udx2m[kc]) * s->usrm[kc] +
                       (QZ(kc, jp, ic) - QZ(kc, jc, ic)) * dxz;
        double a = fabs(dqcap);
        qmax = a > qmax ? a : qmax;
      }
    }
  }
  return qmax;
}

/* ------------------------------------------------------------------ */
/* CalcTorque.F90:26-55   (serial sum, as on one rank)                */
/* ------------------------------------------------------------------ */
void or_calc_torque(or_state *s, double *inner, double *outer) {
  const int nrm = s->nrm, nr = s->nr, nzm = s->nzm, nthm = s->nthm;
  const double rint = s->p.rint, rext = s->p.rext, virotint = s->p.virotint, virotext = s->p.virotext;
  double reninv = 1.0 / s->p.ren;
  double anusslow = 0.0;
  double del = -1.0 / (s->rc[2] - s->rc[1]);
  double anussup = 0.0;
  double deln = -1.0 / (s->rc[nr] - s->rc[nrm]);
  double jlam = 2.0 * rint * rint * rext * rext * (virotint / rint - virotext / rext) /
                (rext * rext - rint * rint) * reninv;
  double jlami = 1.0 / jlam;
  for (int i = 1; i <= nthm; i++)
    for (int j = 1; j <= nzm; j++) {
      double grtlow = ((QT(2, j, i) * s->usrm[2] + QT(1, j, i) * s->usrm[1]) * 0.5 - virotint * s->usrc[1]) * del;
      anusslow = anusslow + grtlow * ICYL(j, i);
      double grtup = (virotext * s->usrc[nr] - (QT(nrm - 1, j, i) * s->usrm[nrm - 1] + QT(nrm, j, i) * s->usrm[nrm]) * 0.5) * deln;
      anussup = anussup + grtup * OCYL(j, i);
    }
  anusslow = anusslow * (s->rc[2] * s->rc[2] * s->rc[2]) * reninv * jlami;
  anussup = anussup * (s->rc[nrm] * s->rc[nrm] * s->rc[nrm]) * reninv * jlami;
  double my_surface = 1.0 / (nthm * nzm);
  *inner = anusslow * my_surface;
  *outer = anussup * my_surface;
}

/* ------------------------------------------------------------------ */
/* CalcVelocityStats.F90:21-60                                        */
/* ------------------------------------------------------------------ */
void or_calc_velocity_stats(or_state *s, double *vmax, double *rms) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double pi = s->pi;
  vmax[0] = vmax[1] = vmax[2] = -100.0;
  double qt2 = 0.0, qr2 = 0.0, qz2 = 0.0;
  double volt = pi * (s->p.rext * s->p.rext - s->p.rint * s->p.rint);
  for (int ic = 1; ic <= nthm; ic++) {
    int ip = ic + 1;
    for (int jc = 1; jc <= nzm; jc++) {
      int jp = jc + 1;
      for (int kc = 1; kc <= nrm; kc++) {
        int kp = kc + 1;
        double a = fabs(QT(kc, jc, ic)), b = fabs(QR(kc, jc, ic)), c = fabs(QZ(kc, jc, ic));
        vmax[0] = vmax[0] > a ? vmax[0] : a;
        vmax[1] = vmax[1] > b ? vmax[1] : b;
        vmax[2] = vmax[2] > c ? vmax[2] : c;
        double t = QT(kc, jc, ic) + QT(kc, jc, ip);
        qt2 = qt2 + t * t * s->pvol[kc] * pi * 0.25;
        double u = QR(kc, jc, ic) * s->usrc[kc] + QR(kp, jc, ic) * s->usrc[kp];
        qr2 = qr2 + u * u * s->pvol[kc] * pi * 0.25;
        double w = QZ(kc, jc, ic) + QZ(kc, jp, ic);
        qz2 = qz2 + w * w * s->pvol[kc] * pi * 0.25;
      }
    }
  }
  double den = (double)nthm * (double)nzm * volt;
  rms[0] = qt2 / den;
  rms[1] = qr2 / den;
  rms[2] = qz2 / den;
}

/* ------------------------------------------------------------------ */
/* Statistics (StatRoutines.F90:12-71) with the vorticity of          */
/* CalcVorticity.F90:11-88 and the dissipation of CalcDissipation.F90 */
/* ------------------------------------------------------------------ */
#define STAT(q, k, j) s->stats[(size_t)(q) * s->nrm * s->nzm + (size_t)((k) - 1) + (size_t)s->nrm * ((j) - 1)]

/* CalcVorticity.F90:30-81 at (kc, jc, ic); halos as update_halo leaves them */
static void vorticity_at(or_state *s, int kc, int jc, int ic, double *vt, double *vr, double *vz) {
  const int im = ic - 1, ip = ic + 1, jm = jc - 1, jp = jc + 1;
  const int kmm = s->kmv[kc], kpp = s->kpv[kc], kp = kc + 1;
  const double dxz = s->dxz, dxt = s->dxt;
  const double *usrc = s->usrc, *usrm = s->usrm, *rm = s->rm, *udx2mm = s->udx2mm;
  double durdz = ((QR(kp, jp, ic) * usrc[kp] + QR(kc, jp, ic) * usrc[kc]) -
                  (QR(kp, jm, ic) * usrc[kp] + QR(kc, jm, ic) * usrc[kc])) * dxz * 0.25;
  double duzdr = ((QZ(kpp, jc, ic) + QZ(kpp, jp, ic)) - (QZ(kmm, jc, ic) + QZ(kmm, jp, ic))) * 0.5 *
                 udx2mm[kc];
  *vt = durdz - duzdr;
  double duzdt = ((QZ(kc, jc, ip) + QZ(kc, jp, ip)) - (QZ(kc, jc, im) + QZ(kc, jp, im))) * 0.25 * dxt *
                 usrm[kc];
  double dutdz = ((QT(kc, jp, ic) + QT(kc, jp, ip)) - (QT(kc, jm, ic) + QT(kc, jm, ip))) * 0.25 * dxz;
  *vr = duzdt - dutdz;
  double dutdr = ((QT(kpp, jc, ic) + QT(kpp, jc, ip)) * rm[kpp] -
                  (QT(kmm, jc, ic) + QT(kmm, jc, ip)) * rm[kmm]) * 0.5 * udx2mm[kc] * usrm[kc];
  double durdt = ((QR(kc, jc, ip) * usrc[kc] + QR(kp, jc, ip) * usrc[kp]) -
                  (QR(kc, jc, im) * usrc[kc] + QR(kp, jc, im) * usrc[kp])) * dxt * 0.25;
  *vz = dutdr - durdt;
}

/* StatRoutines.F90:12-71 (the vortices computed first by CalcVorticity
 * are evaluated at the point: same values).  vtvz_mean keeps the
 * reference's qr(k,j+1,i) in place of qz(k,j+1,i) (StatRoutines.F90:60). */
void or_calc_stats(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double *usrc = s->usrc;
  s->nstatsamples = s->nstatsamples + 1;
  const double usnthm = 1.0 / nthm;
  for (int i = 1; i <= nthm; i++)
    for (int j = 1; j <= nzm; j++)
      for (int k = 1; k <= nrm; k++) {
        const int kp = s->kpv[k];
        double vt, vr, vz;
        vorticity_at(s, k, j, i, &vt, &vr, &vz);
        const double qt = QT(k, j, i), qr = QR(k, j, i), qz = QZ(k, j, i), pr = PR(k, j, i);
        STAT(OR_ST_VT_ME, k, j) += qt * usnthm;
        STAT(OR_ST_VR_ME, k, j) += qr * usnthm * usrc[k];
        STAT(OR_ST_VZ_ME, k, j) += qz * usnthm;
        STAT(OR_ST_WT_ME, k, j) += vt * usnthm;
        STAT(OR_ST_WR_ME, k, j) += vr * usnthm;
        STAT(OR_ST_WZ_ME, k, j) += vz * usnthm;
        STAT(OR_ST_VT_RMS, k, j) += qt * qt * usnthm;
        STAT(OR_ST_VR_RMS, k, j) += (qr * usrc[k]) * (qr * usrc[k]) * usnthm;
        STAT(OR_ST_VZ_RMS, k, j) += qz * qz * usnthm;
        STAT(OR_ST_WT_RMS, k, j) += vt * vt * usnthm;
        STAT(OR_ST_WR_RMS, k, j) += vr * vr * usnthm;
        STAT(OR_ST_WZ_RMS, k, j) += vz * vz * usnthm;
        STAT(OR_ST_PR_ME, k, j) += pr * usnthm;
        STAT(OR_ST_PR_RMS, k, j) += pr * pr * usnthm;
        STAT(OR_ST_VTVR_ME, k, j) += (QT(k, j, i) + QT(k, j, i + 1)) *
                                     (QR(k, j, i) * usrc[k] + QR(kp, j, i) * usrc[kp]) * usnthm * 0.25;
        STAT(OR_ST_VTVZ_ME, k, j) += (QT(k, j, i) + QT(k, j, i + 1)) * (QZ(k, j, i) + QR(k, j + 1, i)) *
                                     usnthm * 0.25;
        STAT(OR_ST_VRVZ_ME, k, j) += (QZ(k, j, i) + QZ(k, j + 1, i)) *
                                     (QR(k, j, i) * usrc[k] + QR(kp, j, i) * usrc[kp]) * usnthm * 0.25;
      }
}

/* CalcDissipation.F90:37-150: accumulates disste and returns nute */
double or_calc_dissipation(or_state *s) {
  const int nrm = s->nrm, nzm = s->nzm, nthm = s->nthm;
  const double dxt = s->dxt, dxz = s->dxz, pi = s->pi, ren = s->p.ren;
  const double *usrc = s->usrc, *usrm = s->usrm, *rm = s->rm, *udx2mm = s->udx2mm,
               *udx2cc = s->udx2cc;
  double my_dissipte = 0.0;
  for (int i = 1; i <= nthm; i++) {
    const int imm = i - 1, ipp = i + 1;
    for (int j = 1; j <= nzm; j++) {
      const int jmm = j - 1, jpp = j + 1;
      for (int k = 1; k <= nrm; k++) {
        const int kp = k + 1, kpp = s->kpv[k], kmm = s->kmv[k];
        double e11 = (QT(k, j, ipp) - QT(k, j, imm)) * dxt * usrm[k] * 0.5 + QR(k, j, i) * (usrc[k] * usrc[k]);
        double e12 = (QT(kpp, j, i) * usrm[kpp] - QT(kmm, j, i) * usrm[kmm]) * udx2mm[k] * rm[k] +
                     (QR(k, j, ipp) - QR(k, j, imm)) * usrc[k] * dxt * 0.5;
        double e13 = (QT(k, jpp, i) - QT(k, jmm, i)) * 0.5 * dxz +
                     (QZ(k, j, ipp) - QZ(k, j, imm)) * 0.5 * dxt * usrm[k];
        double e22 = (QR(kp, j, i) * usrc[kp] - QR(kmm, j, i) * usrc[kmm]) * udx2cc[k];
        double e23 = (QZ(kpp, j, i) - QZ(kmm, j, i)) * udx2mm[k] + (QR(k, jpp, i) - QR(k, jmm, i)) * usrc[k] * dxz * 0.5;
        double e33 = (QZ(k, jpp, i) - QZ(k, jmm, i)) * dxz * 0.5;
        double dissipte = 2.0 * (e11 * e11 + e22 * e22 + e33 * e33) + e12 * e12 + e13 * e13 + e23 * e23;
        my_dissipte = my_dissipte + dissipte * s->pvol[k] * pi;
        STAT(OR_ST_DISSTE, k, j) += dissipte / ren / (double)nthm;
      }
    }
  }
  const double rext = s->p.rext, rint = s->p.rint;
  const double volt = pi * (rext * rext - rint * rint);
  const double w = (s->p.virotint / rint - s->p.virotext / rext) / (rext - rint);
  const double lamdissip = 4.0 * (rext * rext) * (rint * rint) / ((rext + rint) * (rext + rint)) * (w * w);
  s->last_nute = my_dissipte / ((double)nthm * (double)nzm * volt * lamdissip);
  return s->last_nute;
}

int or_get_stats(or_state *s, double *out, int *nsamples) {
  if (out) memcpy(out, s->stats, sizeof(double) * (size_t)OR_NSTATS * s->nrm * s->nzm);
  if (nsamples) *nsamples = s->nstatsamples;
  return 0;
}

double or_last_nute(or_state *s) { return s->last_nute; }

/* ------------------------------------------------------------------ */
/* Restart onto another grid: interp.F90:11-303 as ReadFlowField.F90  */
/* calls it on one rank                                                */
/* ------------------------------------------------------------------ */

/* gridnew (interp.F90:255-303): 1-based rc[1..n], rm[1..n]; ISTR 0 or 6 */
static int gridnew(int n, double rint, double rext, int istr, double str, double *rc, double *rm) {
  if (istr == 0) {
    for (int i = 1; i <= n; i++) {
      double x3 = (double)(i - 1) / (double)(n - 1);
      rc[i] = rext * x3 + rint;
    }
  } else if (istr == 6) {
    const double pi = 2.0 * asin(1.0);
    const int nclip = (int)str, nrmo = n + nclip + nclip;
    double *etazm = (double *)calloc((size_t)nrmo + 2, sizeof(double));
    double *etaz = (double *)calloc((size_t)n + 2, sizeof(double));
    for (int i = 1; i <= nrmo; i++) etazm[i] = +cos(pi * ((double)i - 0.5) / (double)nrmo);
    for (int i = 1; i <= n; i++) etaz[i] = etazm[i + nclip];
    const double delet = etaz[1] - etaz[n];
    for (int i = 1; i <= n; i++) etaz[i] = etaz[i] / (0.5 * delet);
    rc[1] = rint;
    for (int i = 2; i <= n - 1; i++) rc[i] = (rext - rint) * (1. - etaz[i]) * 0.5 + rint;
    rc[n] = rext;
    free(etazm);
    free(etaz);
  } else {
    return -1;
  }
  for (int i = 1; i <= n - 1; i++) rm[i] = (rc[i] + rc[i + 1]) * 0.5;
  rm[n] = 2 * rc[n] - rm[n - 1];
  return 0;
}

/* interp.F90:53-67: second-order extrapolation to indices 0 and n+1 */
static void extrap(double *v, int n) {
  v[0] = 2 * v[1] - v[2];
  v[n + 1] = 2 * v[n] - v[n - 1];
}

/* first l in 1..no with old[l] >= x: (l-1, l); none: (no-1, no) */
static void bracket(const double *old, int no, double x, int *f, int *c) {
  *f = no - 1;
  *c = no;
  for (int l = 1; l <= no; l++)
    if (old[l] >= x) {
      *f = l - 1;
      *c = l;
      return;
    }
}

/* ReadFlowField.F90:63-131 + interp.F90 for one field (intvar 1 qt, 2 qr,
 * 3 qz).  old: (nro, nzo-1, ntho-1) Fortran order; out: (nr, nzm, nthm).
 * The reference leaves the halo edges and corners of its old-grid buffer
 * unset; here they are extrapolated the same way as the faces (radial halo
 * over the whole padded (j, i) range after the theta and z halos). */
int or_regrid(or_state *s, int intvar, const double *old, int ntho, int nzo, int nro, int istro,
              double stro, double *out) {
  const int nth = s->nth, nz = s->nz, nr = s->nr, nthm = s->nthm, nzm = s->nzm;
  if (nth > ntho * 2 || nz > nzo * 2 || nr > nro * 2) {
    snprintf(g_err, sizeof g_err, "new grid more than twice the old resolution");
    return -1;
  }
  const int nthom = ntho - 1, nzom = nzo - 1;
  const int xs2o = (1 * nzom / nzm) > 1 ? (1 * nzom / nzm) : 1;
  const int xs3o = (1 * nthom / nthm) > 1 ? (1 * nthom / nthm) : 1;
  if (xs2o != 1 || xs3o != 1) {
    snprintf(g_err, sizeof g_err, "coarsening by 2 or more in theta or z is not supported");
    return -1;
  }
  const int xe2o = nzom, xe3o = nthom;
  const double pi = s->pi, thmax = 2.0 * pi / s->p.lamb;
  double *tcold = dvec(ntho), *tmold = dvec(ntho), *zcold = dvec(nzo), *zmold = dvec(nzo);
  double *rcold = dvec(nro), *rmold = dvec(nro);
  double *tcnew = dvec(nth), *tmnew = dvec(nth), *zcnew = dvec(nz), *zmnew = dvec(nz);
  double *rcnew = dvec(nr), *rmnew = dvec(nr);
  int rc = 0;
  rc |= gridnew(ntho, 0.0, thmax, 0, 1.0, tcold, tmold);
  rc |= gridnew(nzo, 0.0, s->p.alx3, 0, 1.0, zcold, zmold);
  rc |= gridnew(nro, s->p.rint, s->p.rext, istro, stro, rcold, rmold);
  rc |= gridnew(nth, 0.0, thmax, 0, 1.0, tcnew, tmnew);
  rc |= gridnew(nz, 0.0, s->p.alx3, 0, 1.0, zcnew, zmnew);
  rc |= gridnew(nr, s->p.rint, s->p.rext, s->p.istr, s->p.str, rcnew, rmnew);
  double *bufs[] = {tcold, tmold, zcold, zmold, rcold, rmold, tcnew, tmnew, zcnew, zmnew, rcnew, rmnew};
  if (rc) {
    snprintf(g_err, sizeof g_err, "interp.F90's gridnew builds ISTR 0 and 6 only");
    for (size_t q = 0; q < sizeof bufs / sizeof bufs[0]; q++) free(bufs[q]);
    return -1;
  }
  extrap(tcold, ntho); extrap(tmold, ntho); extrap(zcold, nzo); extrap(zmold, nzo);
  extrap(rcold, nro); extrap(rmold, nro);
  const double *xold, *yold, *zold, *xnew, *ynew, *znew;
  switch (intvar) {
    case 1: xold = tcold; yold = zmold; zold = rmold; xnew = tcnew; ynew = zmnew; znew = rmnew; break;
    case 2: xold = tmold; yold = zcold; zold = rmold; xnew = tmnew; ynew = zcnew; znew = rmnew; break;
    default: xold = tmold; yold = zmold; zold = rcold; xnew = tmnew; ynew = zmnew; znew = rcnew; break;
  }
  /* arrold(0:nro+1, 0:nzo, 0:ntho) */
  const int ak = nro + 2, aj = nzo + 1;
#define AO(k, j, i) a[(size_t)(k) + (size_t)ak * ((size_t)(j) + (size_t)aj * (size_t)(i))]
  double *a = (double *)calloc((size_t)ak * aj * (ntho + 1), sizeof(double));
  for (int i = xs3o; i <= xe3o; i++)
    for (int j = xs2o; j <= xe2o; j++)
      for (int k = 1; k <= nro; k++)
        AO(k, j, i) = old[(size_t)(k - 1) + (size_t)nro * ((size_t)(j - 1) + (size_t)nzom * (i - 1))];
  for (int i = xs3o; i <= xe3o; i++) /* interp.F90:113-118 */
    for (int j = xs2o; j <= xe2o; j++) {
      AO(0, j, i) = 2.0 * AO(1, j, i) - AO(2, j, i);
      AO(nro + 1, j, i) = 2.0 * AO(nro, j, i) - AO(nro - 1, j, i);
    }
  for (int j = xs2o; j <= xe2o; j++) /* :120-125 */
    for (int k = 1; k <= nro; k++) {
      AO(k, j, xs3o - 1) = 2.0 * AO(k, j, xs3o) - AO(k, j, xs3o + 1);
      AO(k, j, xe3o + 1) = 2.0 * AO(k, j, xe3o) - AO(k, j, xe3o - 1);
    }
  for (int i = xs3o; i <= xe3o; i++) /* :127-132 */
    for (int k = 1; k <= nro; k++) {
      AO(k, xs2o - 1, i) = 2.0 * AO(k, xs2o, i) - AO(k, xs2o + 1, i);
      AO(k, xe2o + 1, i) = 2.0 * AO(k, xe2o, i) - AO(k, xe2o - 1, i);
    }
  /* edges and corners the reference leaves unset */
  for (int k = 1; k <= nro; k++)
    for (int ii = 0; ii <= 1; ii++) {
      const int i = ii ? xe3o + 1 : xs3o - 1;
      AO(k, xs2o - 1, i) = 2.0 * AO(k, xs2o, i) - AO(k, xs2o + 1, i);
      AO(k, xe2o + 1, i) = 2.0 * AO(k, xe2o, i) - AO(k, xe2o - 1, i);
    }
  for (int i = 0; i <= ntho; i++)
    for (int j = 0; j <= nzo; j++) {
      if (i >= xs3o && i <= xe3o && j >= xs2o && j <= xe2o) continue;
      AO(0, j, i) = 2.0 * AO(1, j, i) - AO(2, j, i);
      AO(nro + 1, j, i) = 2.0 * AO(nro, j, i) - AO(nro - 1, j, i);
    }
  for (int i = 1; i <= nthm; i++) /* interp.F90:136-215 */
    for (int j = 1; j <= nzm; j++)
      for (int k = 1; k <= nr; k++) {
        const double bix = xnew[i], biy = ynew[j], biz = znew[k];
        int bifi, bici, bifj, bicj, bifk, bick;
        bracket(xold, ntho, bix, &bifi, &bici);
        bracket(yold, nzo, biy, &bifj, &bicj);
        bracket(zold, nro, biz, &bifk, &bick);
        const double bix1 = xold[bifi], bix2 = xold[bici], biy1 = yold[bifj], biy2 = yold[bicj];
        const double biz1 = zold[bifk], biz2 = zold[bick];
        /* every corner value is taken on plane bifk (interp.F90:182-190) */
        const double aficjfk = AO(bifk, bicj, bifi), afifjfk = AO(bifk, bifj, bifi);
        const double acicjfk = AO(bifk, bicj, bici), acifjfk = AO(bifk, bifj, bici);
        const double aficjck = aficjfk, afifjck = afifjfk, acicjck = acicjfk, acifjck = acifjfk;
        const double dxdydz = 1.0 / ((bix2 - bix1) * (biy2 - biy1) * (biz2 - biz1));
        out[(size_t)(k - 1) + (size_t)nr * ((size_t)(j - 1) + (size_t)nzm * (i - 1))] =
            (aficjfk * (bix2 - bix) * (biy - biy1) * (biz2 - biz) +
             afifjfk * (bix2 - bix) * (biy2 - biy) * (biz2 - biz) +
             acicjfk * (bix - bix1) * (biy - biy1) * (biz2 - biz) +
             acifjfk * (bix - bix1) * (biy2 - biy) * (biz2 - biz) +
             aficjck * (bix2 - bix) * (biy - biy1) * (biz - biz1) +
             afifjck * (bix2 - bix) * (biy2 - biy) * (biz - biz1) +
             acicjck * (bix - bix1) * (biy - biy1) * (biz - biz1) +
             acifjck * (bix - bix1) * (biy2 - biy) * (biz - biz1)) * dxdydz;
      }
#undef AO
  free(a);
  for (size_t q = 0; q < sizeof bufs / sizeof bufs[0]; q++) free(bufs[q]);
  return 0;
}

/* DumpRCuts.F90:28-36; outputs (nzm, nthm), Fortran order */
void or_dump_rcut(or_state *s, int kslab, double *vt, double *vr, double *vz) {
  const int nzm = s->nzm;
  for (int i = 1; i <= s->nthm; i++)
    for (int j = 1; j <= nzm; j++) {
      const size_t t = (size_t)(j - 1) + (size_t)nzm * (i - 1);
      vz[t] = QZ(kslab, j, i);
      vt[t] = QT(kslab, j, i);
      vr[t] = (QR(kslab, j, i) * s->usrc[kslab] + QR(kslab + 1, j, i) * s->usrc[kslab + 1]) * 0.5;
    }
}

/* DumpThCuts.F90:29-44; outputs (nrm, nzm), Fortran order */
void or_dump_thcut(or_state *s, double *vt, double *vr, double *vz) {
  const int nrm = s->nrm;
  const size_t m = (size_t)nrm * s->nzm;
  for (size_t t = 0; t < m; t++) vt[t] = vr[t] = vz[t] = 0.0;
  for (int i = 1; i <= s->nthm; i++)
    for (int j = 1; j <= s->nzm; j++)
      for (int k = 1; k <= nrm; k++) {
        if (i == 1) {
          const size_t t = (size_t)(k - 1) + (size_t)nrm * (j - 1);
          vt[t] = vt[t] + QT(k, j, i);
          vr[t] = vr[t] + QR(k, j, i) * s->usrc[k];
          vz[t] = vz[t] + QZ(k, j, i);
        }
      }
}

/* DumpStreamwiseAvg.F90:29-44; outputs (nrm, nzm), Fortran order */
void or_dump_streamwise_avg(or_state *s, double *vt, double *vr, double *vz) {
  const int nrm = s->nrm;
  const size_t m = (size_t)nrm * s->nzm;
  const double usnthm = 1.0 / (double)s->nthm;
  for (size_t t = 0; t < m; t++) vt[t] = vr[t] = vz[t] = 0.0;
  for (int i = 1; i <= s->nthm; i++)
    for (int j = 1; j <= s->nzm; j++)
      for (int k = 1; k <= nrm; k++) {
        const size_t t = (size_t)(k - 1) + (size_t)nrm * (j - 1);
        vt[t] = vt[t] + QT(k, j, i) * usnthm;
        vr[t] = vr[t] + QR(k, j, i) * s->usrc[k] * usnthm;
        vz[t] = vz[t] + QZ(k, j, i) * usnthm;
      }
}

/* STREAM triad a = b + s*c on the host cores (OpenMP): the host's
 * achievable bandwidth, reported beside the CPU baseline.  Returns GB/s of
 * the best of reps passes over n doubles per array. */
double or_stream_triad(long long n, int reps) {
  double *a = (double *)malloc(sizeof(double) * (size_t)n);
  double *b = (double *)malloc(sizeof(double) * (size_t)n);
  double *c = (double *)malloc(sizeof(double) * (size_t)n);
  if (!a || !b || !c) {
    free(a); free(b); free(c);
    return 0.0;
  }
#pragma omp parallel for schedule(static)
  for (long long q = 0; q < n; q++) {
    a[q] = 0.0;
    b[q] = 1.0;
    c[q] = 2.0;
  }
  double best = 1e30;
  for (int r = 0; r < reps; r++) {
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < n; q++) a[q] = b[q] + 3.0 * c[q];
#ifdef _OPENMP
    const double t = omp_get_wtime() - t0;
    if (t < best) best = t;
#endif
  }
  const double sum = a[n / 2];
  free(a); free(b); free(c);
  return sum > 0.0 && best < 1e30 ? 24.0 * (double)n / best / 1e9 : 0.0;
}

/* ------------------------------------------------------------------ */
/* main.F90:152-223: one iteration of the step loop                    */
/* ------------------------------------------------------------------ */
int or_step(or_state *s, int ntime, double walltime, double *o) {
  int errorcode = 0;
  double instcfl = or_calc_max_cfl(s);
  if (s->variabletstep) {
    if (ntime > 1) {
      if (instcfl < 1.0e-8)
        s->dt = s->p.dtmax;
      else
        s->dt = s->p.limitcfl / instcfl;
      if (s->dt > s->p.dtmax) s->dt = s->p.dtmax;
    } else {
      s->dt = s->p.dtmin;
    }
    if (s->dt < s->p.dtmin) errorcode = 166;
  } else {
    s->dt = s->p.dtmax; /* intended fixed dt (main.F90:63-65) */
    instcfl = instcfl * s->dt;
    if (instcfl > s->p.limitcfl) errorcode = 165;
  }
  o[0] = instcfl;
  o[1] = s->dt;
  or_time_marcher(s);
  s->time = s->time + s->dt;
  o[2] = s->time;
  o[3] = o[4] = o[5] = o[6] = o[7] = o[8] = 0.0;
  if (ntime == 1 || fmod(s->time, s->p.tout) < s->dt) {
    double vmax[3], rms[3];
    or_calc_velocity_stats(s, vmax, rms);
    if (vmax[0] > s->p.limitvel && vmax[1] > s->p.limitvel) errorcode = 266;
    o[0] = or_calc_max_cfl(s);
    double dmax = or_check_divergence(s);
    or_calc_torque(s, &o[4], &o[5]);
    o[3] = dmax;
    o[6] = vmax[0];
    o[7] = vmax[1];
    o[8] = vmax[2];
    if (fabs(dmax) > s->p.resid) errorcode = 169;
    if (s->time > s->p.tsta) { /* main.F90:188-190 */
      or_calc_stats(s);
      or_calc_dissipation(s);
    }
    if (!s->variabletstep) o[0] = o[0] * s->dt; /* main.F90:201 */
  }
  if (s->time > s->p.tmax) errorcode = 333;
  if (walltime > s->p.walltimemax) errorcode = 334;
  if (ntime == s->p.ntst) errorcode = 555;
  return errorcode;
}
