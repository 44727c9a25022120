// Host-side grid, metric and coefficient vectors of the Taylor-Couette solver.
// Restates CreateGrid.F90:24-323, InitPressureSolver.F90:22-65,
// InitTimeMarchScheme.F90:16-48, SlipBCs.F90:31-90 and InnerCylVelocity.F90:15-16.
#pragma once
#include <string>
#include <vector>

#include "../../include/tc_abi.h"

namespace tc {

// 1-based vectors: v[k] is the reference's v(k); index 0 unused.
struct Grid {
  int nthm = 0, nzm = 0, nrm = 0, nth = 0, nz = 0, nr = 0;
  double pi = 0, dxt = 0, dxz = 0, dxr = 0, dxtq = 0, dxzq = 0, dxrq = 0;
  std::vector<double> tc, tm, zc, zm, rc, rm, g2rc, g2rm, g2rmm, g2rcc, pvol;
  std::vector<double> usrm, usrc, udx1vis1, udx1vis2, udx1qm, udx1qc, udx2c, udx2m,
      udx2cc, udx2mm, rudx2c, udx2rm;
  std::vector<int> kmc, kpc, kmv, kpv;
  std::vector<double> ap1j, am1j, ac1j, ap2j, am2j, ac2j, ap2je, am2je, ac2je, ap3j,
      am3j, ac3j, apph, amph, acph;
  double ap1j_fs[3] = {0, 0, 0}, ac1j_fs[3] = {0, 0, 0}, am1j_fs[3] = {0, 0, 0};
  double ap3j_fs[3] = {0, 0, 0}, ac3j_fs[3] = {0, 0, 0}, am3j_fs[3] = {0, 0, 0};
  std::vector<double> ak1, ak3;  // ak1(1..nthm), ak3(1..nzm)
  double gam[4] = {0, 0, 0, 0}, rom[4] = {0, 0, 0, 0}, alm[4] = {0, 0, 0, 0};
};

// Validates p and fills g; returns an empty string on success or the reason.
std::string build_grid(const tc_params &p, Grid &g);

// SlipBCs.F90:31-90 at time t: masks (nzm,nthm) column-major.
void slip_bcs(const tc_params &p, const Grid &g, double time, std::vector<int> &icyl,
              std::vector<int> &ocyl);

// Returns the requested vector (Fortran 1..n) in out; -1 if id is unknown.
int grid_vector(const Grid &g, int id, double *out, int n);

// Partial-pivoting tridiagonal factorisation and solve with the reference
// LAPACK semantics (dgttrf / dgtts2), 0-based.
int dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv);

}  // namespace tc
