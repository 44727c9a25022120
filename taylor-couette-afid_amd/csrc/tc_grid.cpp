// Host-side grid and coefficient construction (runs once per context).
// Each block cites the reference lines whose formulas it evaluates; the
// statement order is kept so the vectors are bit-identical to the reference's.
#include "tc_grid.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace tc {

static void init_time_march(const tc_params &p, Grid &g) {
  // InitTimeMarchScheme.F90:16-48
  if (p.nsst > 1) {
    g.gam[1] = 8.0 / 15.0; g.gam[2] = 5.0 / 12.0; g.gam[3] = 3.0 / 4.0;
    g.rom[1] = 0.0; g.rom[2] = -17.0 / 60.0; g.rom[3] = -5.0 / 12.0;
  } else {
    g.gam[1] = 1.5; g.gam[2] = 0.0; g.gam[3] = 0.0;
    g.rom[1] = -0.5; g.rom[2] = 0.0; g.rom[3] = 0.0;
  }
  for (int ns = 1; ns <= p.nsst; ns++) g.alm[ns] = g.gam[ns] + g.rom[ns];
}

std::string build_grid(const tc_params &p, Grid &g) {
  if (p.nthm < 2 || p.nzm < 2 || p.nrm < 3) return "grid too small (need NTM>=2, NZM>=2, NRM>=3)";
  // the radial Poisson solve holds at most 26 rows per thread, 512 threads per line
  if (p.nrm > 26 * 512) return "grid too large (need NRM <= 13312)";
  if (p.nsst != 1 && p.nsst != 3) return "NSST must be 1 (AB2) or 3 (RK3)";
  if (p.istr != 0 && p.istr != 4 && p.istr != 6) return "ISTR must be 0, 4 or 6";
  if (p.ren <= 0) return "1/NU must be positive";
  if (p.rext <= p.rint || p.rint <= 0) return "need 0 < RINT < REXT";
  if ((p.minstrt * p.multmin) != 0 && p.minstrz == 0)
    return "MINSTRT*MULTMIN != 0 with MINSTRZ == 0 divides by zero in SlipBCs";
  g.nthm = p.nthm; g.nzm = p.nzm; g.nrm = p.nrm;
  g.nth = p.nthm + 1; g.nz = p.nzm + 1; g.nr = p.nrm + 1;
  const int nthm = g.nthm, nzm = g.nzm, nrm = g.nrm, nr = g.nr;
  g.pi = 2.0 * std::asin(1.0);
  const double pi = g.pi;
  init_time_march(p, g);

  auto vec = [](int n) { return std::vector<double>((size_t)n + 2, 0.0); };
  g.tc = vec(g.nth); g.tm = vec(g.nth); g.zc = vec(g.nz); g.zm = vec(g.nz);
  g.rc = vec(nr); g.rm = vec(nr); g.g2rc = vec(nr); g.g2rm = vec(nr); g.g2rmm = vec(nr);
  g.g2rcc = vec(nr); g.pvol = vec(nr); g.usrm = vec(nr); g.usrc = vec(nr);
  g.udx1vis1 = vec(nr); g.udx1vis2 = vec(nr); g.udx1qm = vec(nr); g.udx1qc = vec(nr);
  g.udx2c = vec(nr); g.udx2m = vec(nr); g.udx2cc = vec(nr); g.udx2mm = vec(nr);
  g.rudx2c = vec(nr); g.udx2rm = vec(nr);
  g.kmc.assign(nr + 2, 0); g.kpc.assign(nr + 2, 0); g.kmv.assign(nr + 2, 0); g.kpv.assign(nr + 2, 0);
  g.ap1j = vec(nr); g.am1j = vec(nr); g.ac1j = vec(nr); g.ap2j = vec(nr); g.am2j = vec(nr);
  g.ac2j = vec(nr); g.ap2je = vec(nr); g.am2je = vec(nr); g.ac2je = vec(nr); g.ap3j = vec(nr);
  g.am3j = vec(nr); g.ac3j = vec(nr); g.apph = vec(nr); g.amph = vec(nr); g.acph = vec(nr);
  g.ak1 = vec(g.nth); g.ak3 = vec(g.nz);

  // CreateGrid.F90:28-34
  g.dxt = ((double)nthm * p.lamb) / (2. * pi);
  g.dxz = (double)nzm / p.alx3;
  g.dxr = (double)nrm / 1.0;
  g.dxtq = g.dxt * g.dxt;
  g.dxzq = g.dxz * g.dxz;
  g.dxrq = g.dxr * g.dxr;
  const double dxt = g.dxt, dxr = g.dxr, dxtq = g.dxtq, dxrq = g.dxrq;

  // CreateGrid.F90:40-52: wall-clamped radial neighbour indices
  for (int kc = 1; kc <= nrm; kc++) {
    g.kmv[kc] = (kc == 1) ? kc : kc - 1;
    g.kpv[kc] = (kc == nrm) ? kc : kc + 1;
    g.kpc[kc] = g.kpv[kc] - kc;
    g.kmc[kc] = kc - g.kmv[kc];
  }
  // CreateGrid.F90:57-73: uniform theta and z
  for (int i = 1; i <= g.nth; i++) g.tc[i] = (double)(i - 1) / dxt;
  for (int i = 1; i <= nthm; i++) g.tm[i] = ((double)(i - 1) + 0.5) / dxt;
  for (int j = 1; j <= g.nz; j++) g.zc[j] = (double)(j - 1) / g.dxz;
  for (int j = 1; j <= nzm; j++) g.zm[j] = ((double)(j - 1) + 0.5) / g.dxz;

  // CreateGrid.F90:83-128: radial nodes
  std::vector<double> &rc = g.rc, &rm = g.rm, &g2rm = g.g2rm, &g2rc = g.g2rc;
  const double rint = p.rint, rext = p.rext, str = p.str;
  if (p.istr == 0) {
    for (int k = 1; k <= nr; k++) {
      double x3 = (double)(k - 1) / (double)nrm;
      rc[k] = (rext - rint) * x3 + rint;
    }
  } else if (p.istr == 4) {
    const double tstr = std::tanh(str);
    rc[1] = rint;
    for (int k = 2; k <= nr; k++) {
      double z2dp = (double)(2 * k - nr - 1) / (double)nrm;
      rc[k] = (1 + std::tanh(str * z2dp) / tstr) * 0.5 * (rext - rint) + rint;
    }
  } else {
    const int nclip = (int)str;
    const int nrmo = nr + nclip + nclip;
    std::vector<double> etazm((size_t)nrmo + 2, 0.0), etaz((size_t)nr + 2, 0.0);
    for (int k = 1; k <= nrmo; k++) etazm[k] = +std::cos(pi * ((double)k - 0.5) / (double)nrmo);
    for (int k = 1; k <= nr; k++) etaz[k] = etazm[k + nclip];
    const double delet = etaz[1] - etaz[nr];
    for (int k = 1; k <= nr; k++) etaz[k] = etaz[k] / (0.5 * delet);
    rc[1] = rint;
    for (int k = 2; k <= nrm; k++) rc[k] = (rext - rint) * (1. - etaz[k]) * 0.5 + rint;
    rc[nr] = rext;
  }
  for (int k = 2; k <= nr; k++)
    if (!(rc[k] > rc[k - 1])) return "radial grid is not increasing (check STR3)";

  // CreateGrid.F90:136-174: staggered nodes and metrics
  for (int k = 1; k <= nrm; k++) {
    rm[k] = (rc[k] + rc[k + 1]) * 0.5;
    g2rm[k] = (rc[k + 1] - rc[k]) * dxr;
  }
  g2rc[1] = (rc[2] - rc[1]) * dxr;
  g.g2rmm[1] = (rm[2] - rm[1]) * dxr;
  g.g2rcc[1] = (rc[2] - rc[1]) * dxr;
  for (int k = 2; k <= nrm; k++) {
    g2rc[k] = (rc[k + 1] - rc[k - 1]) * dxr * 0.5;
    g.g2rmm[k] = (rm[k + 1] - rm[k - 1]) * dxr;
    g.g2rcc[k] = (rc[k + 1] - rc[k - 1]) * dxr;
  }
  g2rc[nr] = (rc[nr] - rc[nrm]) * dxr;
  g.g2rmm[nr] = (rm[nr] - rm[nrm]) * dxr;
  g.g2rcc[nr] = (rc[nr] - rc[nrm]) * dxr;
  for (int k = 1; k <= nrm; k++) {
    g.udx1vis1[k] = 2.0 * dxt / (rm[k] * rm[k] * p.ren);
    g.udx2rm[k] = dxr / (g2rm[k] * rm[k]);
    g.udx2m[k] = dxr / g2rm[k];
    g.udx2mm[k] = dxr / g.g2rmm[k];
    g.udx2c[k] = dxr / g2rc[k];
    g.rudx2c[k] = g.udx2c[k] * rc[k];
    g.usrm[k] = 1.0 / rm[k];
    g.udx1qm[k] = dxtq / (rm[k] * rm[k]);
    g.pvol[k] = (rc[k + 1] * rc[k + 1] - rc[k] * rc[k]);
  }
  g.udx2c[nr] = dxr / g2rc[nr];
  for (int k = 1; k <= nr; k++) {
    g.udx1vis2[k] = dxt / (rc[k] * p.ren);
    g.usrc[k] = 1.0 / rc[k];
    g.udx1qc[k] = dxtq / (rc[k] * rc[k]);
    g.udx2cc[k] = dxr / g.g2rcc[k];
  }

  // q_theta operator, CreateGrid.F90:194-238
  for (int kc = 2; kc <= nrm - 1; kc++) {
    const int kp = kc + 1;
    const double a22 = dxrq / g2rm[kc] / rm[kc];
    const double a22p = +a22 * rc[kp] / g2rc[kp];
    const double a22m = +a22 * rc[kc] / g2rc[kc];
    g.ap1j[kc] = a22p;
    g.am1j[kc] = a22m;
    g.ac1j[kc] = -(a22p + a22m) - 1. / (rm[kc] * rm[kc]);
  }
  {
    const int kc = nrm, kp = kc + 1;
    const double a22 = dxrq / (g2rm[kc] * rm[kc]);
    const double apjjc = a22 * rc[kp] / g2rc[kp];
    const double amjjc = a22 * rc[kc] / g2rc[kc];
    g.ap1j[kc] = 2. * apjjc;
    g.ac1j[kc] = -2. * apjjc - amjjc - 1. / (rm[kc] * rm[kc]);
    g.am1j[kc] = amjjc;
    // the outer-wall free-slip row lands in slot 1 and is overwritten by the
    // inner-wall one just below; slot 2 keeps 0 (CreateGrid.F90:218-220)
    g.ap1j_fs[1] = 0.;
    g.ac1j_fs[1] = -amjjc;
    g.am1j_fs[1] = amjjc;
  }
  {
    const int kc = 1, kp = kc + 1;
    const double urpp = 1. / (g2rm[kc] * g2rc[kp]);
    const double a22 = dxrq / g2rm[kc] / rm[kc];
    const double a22p = +a22 * rc[kp] / g2rc[kp];
    const double a22m = +a22 * rc[kc] / g2rc[kc];
    g.ap1j[kc] = a22p;
    g.am1j[kc] = 2. * a22m;
    g.ac1j[kc] = -(a22p + 2. * a22m) - 1. / (rm[kc] * rm[kc]);
    g.ap1j_fs[1] = rc[kp] * rc[kp] * rc[kp] / (rm[kc] * rm[kc] * rm[kp]) * dxrq * urpp;
    g.ac1j_fs[1] = -rc[kp] * rc[kp] * rc[kp] / (rm[kc] * rm[kc] * rm[kc]) * dxrq * urpp;
    g.am1j_fs[1] = 0.;
  }
  // q_r implicit operator, CreateGrid.F90:244-259
  g.am2j[1] = 0.; g.ap2j[1] = 0.; g.ac2j[1] = 1.;
  for (int kc = 2; kc <= nrm; kc++) {
    const int km = kc - 1;
    const double a22 = rc[kc] * dxrq / g2rc[kc];
    const double a22p = 1. / (rm[kc] * g2rm[kc]);
    const double a22m = 1. / (rm[km] * g2rm[km]);
    g.ap2j[kc] = a22 * a22p;
    g.am2j[kc] = a22 * a22m;
    g.ac2j[kc] = -(a22 * a22p + a22 * a22m);
  }
  g.am2j[nr] = 0.; g.ap2j[nr] = 0.; g.ac2j[nr] = 1.;
  // q_r explicit operator, CreateGrid.F90:266-276
  for (int kc = 2; kc <= nrm; kc++) {
    const int km = kc - 1, kp = kc + 1;
    const double ugmm = dxrq / g2rc[kc];
    const double a22p = rm[kc] / (rc[kp] * g2rm[kc]);
    const double a22m = rm[km] / (rc[km] * g2rm[km]);
    const double a22 = (rm[kc] / g2rm[kc] + rm[km] / g2rm[km]) / rc[kc];
    g.ap2je[kc] = a22p * ugmm;
    g.am2je[kc] = a22m * ugmm;
    g.ac2je[kc] = -a22 * ugmm - 1. / (rc[kc] * rc[kc]);
  }
  // q_z operator, CreateGrid.F90:281-322
  for (int kc = 2; kc <= nrm - 1; kc++) {
    const int kp = kc + 1;
    const double a22 = dxrq / g2rm[kc] / rm[kc];
    const double a22p = +a22 * rc[kp] / g2rc[kp];
    const double a22m = +a22 * rc[kc] / g2rc[kc];
    g.ap3j[kc] = a22p;
    g.am3j[kc] = a22m;
    g.ac3j[kc] = -(a22p + a22m);
  }
  {
    const int kc = 1, kp = kc + 1;
    const double a22 = dxrq / g2rm[kc] / rm[kc];
    const double a22p = +a22 * rc[kp] / g2rc[kp];
    const double a22m = +a22 * rc[kc] / g2rc[kc];
    g.am3j[kc] = 0.;
    g.ap3j[kc] = a22p;
    g.ac3j[kc] = -(a22p + 2. * a22m);
    g.am3j_fs[1] = 0.;
    g.ap3j_fs[1] = a22p;
    g.ac3j_fs[1] = -(a22p);
  }
  {
    const int kc = nrm, kp = kc + 1;
    const double a22 = dxrq / g2rm[kc] / rm[kc];
    const double a22m = +a22 * rc[kc] / g2rc[kc];
    const double a22p = +a22 * rc[kp] / g2rc[kp];
    g.am3j[kc] = a22m;
    g.ap3j[kc] = 0.;
    g.ac3j[kc] = -(a22m + 2. * a22p);
    g.am3j_fs[2] = a22m;
    g.ap3j_fs[2] = 0.;
    g.ac3j_fs[2] = -(a22m);
  }

  // Poisson operator and modified wavenumbers, InitPressureSolver.F90:22-65
  for (int kc = 1; kc <= nrm; kc++) {
    const int kp = g.kpv[kc];
    const double a22icc = rc[kc] * g.kmc[kc] * dxrq / g2rc[kc];
    const double a22icp = rc[kp] * g.kpc[kc] * dxrq / g2rc[kp];
    const double ac2 = -(a22icc + a22icp);
    const double ugmmm = 1. / (rm[kc] * g2rm[kc]);
    g.amph[kc] = a22icc * ugmmm;
    g.apph[kc] = a22icp * ugmmm;
    g.acph[kc] = ac2 * ugmmm;
  }
  {
    const int nzmh = nzm / 2 + 1, nthmh = nthm / 2 + 1;
    std::vector<double> ao((size_t)nzm + 2), ap((size_t)nthm + 2);
    for (int j = 1; j <= nzmh; j++) ao[j] = (j - 1) * 2. * pi;
    for (int j = nzmh + 1; j <= nzm; j++) ao[j] = -((nzm - j + 1) * 2. * pi);
    const double t = (double)nzm / p.alx3;
    for (int j = 1; j <= nzm; j++) g.ak3[j] = 2. * (1. - std::cos(ao[j] / nzm)) * (t * t);
    for (int i = 1; i <= nthmh; i++) ap[i] = (i - 1) * 2.0 * pi;
    for (int i = nthmh + 1; i <= nthm; i++) ap[i] = -((nthm - i + 1) * 2.0 * pi);
    for (int i = 1; i <= nthm; i++) g.ak1[i] = 2. * (1. - std::cos(ap[i] / nthm)) * dxtq;
  }
  return std::string();
}

void slip_bcs(const tc_params &p, const Grid &g, double time, std::vector<int> &icyl,
              std::vector<int> &ocyl) {
  // SlipBCs.F90:31-90; masks are 1 everywhere unless MINSTR*·MULTMIN != 0
  const int nthm = g.nthm, nzm = g.nzm;
  icyl.assign((size_t)nzm * nthm, 1);
  ocyl.assign((size_t)nzm * nthm, 1);
  const int ntott = p.minstrt * p.multmin, ntotz = p.minstrz * p.multmin;
  const int ntot = ntott + ntotz;
  if (ntot == 0) return;
  const int acute = 1;
  const double perdist = 2 * g.pi / p.lamb;
  const double wint = p.virotint / p.rint;
  double anglemoved = std::fmod(wint * time, perdist);
  const int points_moved = (int)std::floor(anglemoved * g.dxt);
  std::vector<int> thdispl((size_t)nthm + 1);
  for (int i = 1; i <= nthm; i++) {
    thdispl[i] = (i + points_moved) % nthm;
    if (thdispl[i] == 0) thdispl[i] = nthm;
  }
  auto at = [&](std::vector<int> &m, int j, int i) -> int & { return m[(size_t)(j - 1) + (size_t)nzm * (i - 1)]; };
  std::fill(icyl.begin(), icyl.end(), 0);
  const double m1 = (double)p.minstrt / (double)p.minstrz * (double)nzm / (double)nthm;
  const double y1 = (double)nzm / (double)ntotz;
  const double x1 = (double)nthm / (double)ntott;
  int limit1 = 0, limit2 = ntot;
  if (acute == 1 && ntotz != 0) {
    limit1 = -ntott;
    limit2 = ntotz;
  }
  for (int ii = limit1; ii <= limit2; ii++)
    for (int j = 1; j <= nthm; j++) {
      const int truej = thdispl[j];
      if (ntotz == 0 && (j - ii * x1 > 0.00001) && (j - ii * x1 - x1 / 2. < 0.0001)) {
        for (int k = 1; k <= nzm; k++) at(icyl, k, truej) = 1;
      } else if (ntotz != 0) {
        for (int k = 1; k <= nzm; k++)
          if ((k - acute * m1 * j - ii * y1 > 0.00001) && (k - acute * m1 * j - ii * y1 - y1 / 2. < 0.00001))
            at(icyl, k, truej) = 1;
      }
    }
}

int grid_vector(const Grid &g, int id, double *out, int n) {
  const std::vector<double> *v = nullptr;
  int len = g.nr;
  switch (id) {
    case TC_V_RC: v = &g.rc; break;
    case TC_V_RM: v = &g.rm; break;
    case TC_V_G2RC: v = &g.g2rc; break;
    case TC_V_G2RM: v = &g.g2rm; break;
    case TC_V_USRC: v = &g.usrc; break;
    case TC_V_USRM: v = &g.usrm; break;
    case TC_V_UDX2RM: v = &g.udx2rm; break;
    case TC_V_UDX2M: v = &g.udx2m; break;
    case TC_V_UDX2C: v = &g.udx2c; break;
    case TC_V_RUDX2C: v = &g.rudx2c; break;
    case TC_V_UDX1VIS1: v = &g.udx1vis1; break;
    case TC_V_UDX1VIS2: v = &g.udx1vis2; break;
    case TC_V_UDX1QM: v = &g.udx1qm; break;
    case TC_V_AP1J: v = &g.ap1j; break;
    case TC_V_AM1J: v = &g.am1j; break;
    case TC_V_AC1J: v = &g.ac1j; break;
    case TC_V_AP2J: v = &g.ap2j; break;
    case TC_V_AM2J: v = &g.am2j; break;
    case TC_V_AC2J: v = &g.ac2j; break;
    case TC_V_AP2JE: v = &g.ap2je; break;
    case TC_V_AM2JE: v = &g.am2je; break;
    case TC_V_AC2JE: v = &g.ac2je; break;
    case TC_V_AP3J: v = &g.ap3j; break;
    case TC_V_AM3J: v = &g.am3j; break;
    case TC_V_AC3J: v = &g.ac3j; break;
    case TC_V_APPH: v = &g.apph; break;
    case TC_V_AMPH: v = &g.amph; break;
    case TC_V_ACPH: v = &g.acph; break;
    case TC_V_AK1: v = &g.ak1; len = g.nthm; break;
    case TC_V_AK3: v = &g.ak3; len = g.nzm; break;
    case TC_V_PVOL: v = &g.pvol; break;
    case TC_V_FS: {
      const double t[15] = {g.ap1j_fs[1], g.ap1j_fs[2], g.ac1j_fs[1], g.ac1j_fs[2],
                            g.am1j_fs[1], g.am1j_fs[2], g.ap3j_fs[1], g.ap3j_fs[2],
                            g.ac3j_fs[1], g.ac3j_fs[2], g.am3j_fs[1], g.am3j_fs[2],
                            g.dxt, g.dxz, g.dxr};
      const int m = n < 15 ? n : 15;
      std::memcpy(out, t, sizeof(double) * (size_t)m);
      return m;
    }
    default: return -1;
  }
  const int m = n < len ? n : len;
  for (int q = 0; q < m; q++) out[q] = (*v)[(size_t)q + 1];
  return m;
}

int dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv) {
  // LAPACK dgttrf: elimination with row interchange when |dl(i)| > |d(i)|
  for (int i = 0; i < n; i++) ipiv[i] = i;
  for (int i = 0; i < n - 2; i++) du2[i] = 0.0;
  for (int i = 0; i < n - 1; i++) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (d[i] != 0.0) {
        const double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      const double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      const double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      if (i < n - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -fact * du[i + 1];
      }
      ipiv[i] = i + 1;
    }
  }
  for (int i = 0; i < n; i++)
    if (d[i] == 0.0) return i + 1;
  return 0;
}

}  // namespace tc
