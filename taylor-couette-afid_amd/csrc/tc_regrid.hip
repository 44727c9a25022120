// Restart onto another grid (ReadFlowField.F90:63-131 with interp.F90):
// the fields of a continuation file written on an old grid are interpolated
// onto this context's grid.  The 1-D coordinate brackets are host work; the
// gather-and-weight of every new point runs on the device.
//
// Kept as the reference computes it:
//   - coordinates come from interp.F90's own gridnew (ISTR 0 or 6 only; its
//     ISTR 0 radial formula is rext*x + rint), for the old and the new grid;
//   - the staggering interp.F90 assigns: qt on (tc, zm, rm), qr on
//     (tm, zc, rm), qz on (tm, zm, rc);
//   - each new point takes the first old coordinate >= it as the upper
//     bracket, and all eight corner values from the lower radial plane, so
//     the result is bilinear in theta and z and piecewise constant in r;
//   - the old buffer's faces are extrapolated linearly.  Its edges and
//     corners, which the reference leaves unset, are extrapolated the same
//     way.
#include "tc_internal.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace tc {

namespace {

struct Bracket {
  int f, c;           // lower and upper old index (0 .. no)
  double x1, x2, x;   // their coordinates and the new one
};

// interp.F90:255-303: rc, rm at 1..n (index 0 and n+1 left for extrapolation)
bool gridnew(int n, double rint, double rext, int istr, double str, std::vector<double> &rc,
             std::vector<double> &rm) {
  rc.assign((size_t)n + 2, 0.0);
  rm.assign((size_t)n + 2, 0.0);
  if (istr == 0) {
    for (int i = 1; i <= n; i++) rc[i] = rext * ((double)(i - 1) / (double)(n - 1)) + rint;
  } else if (istr == 6) {
    const double pi = 2.0 * std::asin(1.0);
    const int nclip = (int)str, nrmo = n + 2 * nclip;
    std::vector<double> eta((size_t)n + 2);
    for (int i = 1; i <= n; i++) eta[i] = std::cos(pi * ((double)(i + nclip) - 0.5) / (double)nrmo);
    const double delet = eta[1] - eta[n];
    for (int i = 1; i <= n; i++) eta[i] = eta[i] / (0.5 * delet);
    rc[1] = rint;
    for (int i = 2; i <= n - 1; i++) rc[i] = (rext - rint) * (1. - eta[i]) * 0.5 + rint;
    rc[n] = rext;
  } else {
    return false;
  }
  for (int i = 1; i <= n - 1; i++) rm[i] = (rc[i] + rc[i + 1]) * 0.5;
  rm[n] = 2 * rc[n] - rm[n - 1];
  return true;
}

void extrapolate_ends(std::vector<double> &v, int n) {
  v[0] = 2 * v[1] - v[2];
  v[(size_t)n + 1] = 2 * v[n] - v[(size_t)n - 1];
}

// interp.F90:142-176 for the new coordinates xnew[1..m]
std::vector<Bracket> brackets(const std::vector<double> &old, int no, const std::vector<double> &xnew,
                              int m) {
  std::vector<Bracket> b((size_t)m);
  for (int i = 1; i <= m; i++) {
    const double x = xnew[i];
    int f = no - 1, c = no;
    for (int l = 1; l <= no; l++)
      if (old[l] >= x) {
        f = l - 1;
        c = l;
        break;
      }
    b[(size_t)i - 1] = Bracket{f, c, old[f], old[c], x};
  }
  return b;
}

// One new point per thread, device layout [k][j][i]; the old buffer is
// (0:nro+1, 0:nzo, 0:ntho), r fastest.
__global__ __launch_bounds__(256) void k_regrid(const double *__restrict__ a, int ak, int aj,
                                                const Bracket *__restrict__ bt,
                                                const Bracket *__restrict__ bz,
                                                const Bracket *__restrict__ br, int nthm, int nzm,
                                                int nr, double *__restrict__ out) {
  const long long P = (long long)nzm * nthm;
  const long long l = (long long)blockIdx.x * 256 + threadIdx.x;
  if (l >= P * nr) return;
  const int k = (int)(l / P);
  const long long rem = l - (long long)k * P;
  const int j = (int)(rem / nthm), i = (int)(rem - (long long)j * nthm);
  const Bracket X = bt[i], Y = bz[j], Z = br[k];
  auto ao = [&](int kk, int jj, int ii) {
    return a[(long long)kk + (long long)ak * ((long long)jj + (long long)aj * ii)];
  };
  const double aficjfk = ao(Z.f, Y.c, X.f), afifjfk = ao(Z.f, Y.f, X.f);
  const double acicjfk = ao(Z.f, Y.c, X.c), acifjfk = ao(Z.f, Y.f, X.c);
  const double bix = X.x, bix1 = X.x1, bix2 = X.x2, biy = Y.x, biy1 = Y.x1, biy2 = Y.x2;
  const double biz = Z.x, biz1 = Z.x1, biz2 = Z.x2;
  const double dxdydz = 1.0 / ((bix2 - bix1) * (biy2 - biy1) * (biz2 - biz1));
  // interptrilin (interp.F90:218-253) with the upper-plane values equal to
  // the lower-plane ones
  out[l] = (aficjfk * (bix2 - bix) * (biy - biy1) * (biz2 - biz) +
            afifjfk * (bix2 - bix) * (biy2 - biy) * (biz2 - biz) +
            acicjfk * (bix - bix1) * (biy - biy1) * (biz2 - biz) +
            acifjfk * (bix - bix1) * (biy2 - biy) * (biz2 - biz) +
            aficjfk * (bix2 - bix) * (biy - biy1) * (biz - biz1) +
            afifjfk * (bix2 - bix) * (biy2 - biy) * (biz - biz1) +
            acicjfk * (bix - bix1) * (biy - biy1) * (biz - biz1) +
            acifjfk * (bix - bix1) * (biy2 - biy) * (biz - biz1)) * dxdydz;
}

}  // namespace

int regrid_field(tc_ctx *c, int field, const double *old, int ntho, int nzo, int nro, int istro,
                 double stro) {
  if (field != TC_QT && field != TC_QR && field != TC_QZ)
    return set_error("tc_regrid_field: field must be TC_QT, TC_QR or TC_QZ");
  if (!old) return set_error("tc_regrid_field: null array");
  const int nth = c->nthm + 1, nz = c->nzm + 1, nr = c->nr;
  if (ntho < 3 || nzo < 3 || nro < 3) return set_error("tc_regrid_field: old grid too small");
  // ReadFlowField.F90:67-76
  if (nth > ntho * 2 || nz > nzo * 2 || nr > nro * 2)
    return set_error("New grid resolution cannot be more than twice the old resolution");
  // ReadFlowField.F90:84-92 on one rank: the old rows start at
  // max(1, nzom/nzm); past 1 the reference reads outside its buffer
  const int nthom = ntho - 1, nzom = nzo - 1;
  if (nzom / c->nzm > 1 || nthom / c->nthm > 1)
    return set_error("tc_regrid_field: coarsening by 2 or more in theta or z is not supported");
  const tc_params &p = c->p;
  const double thmax = 2.0 * c->g.pi / p.lamb;
  std::vector<double> tco, tmo, zco, zmo, rco, rmo, tcn, tmn, zcn, zmn, rcn, rmn;
  bool ok = gridnew(ntho, 0.0, thmax, 0, 1.0, tco, tmo) && gridnew(nzo, 0.0, p.alx3, 0, 1.0, zco, zmo) &&
            gridnew(nro, p.rint, p.rext, istro, stro, rco, rmo) &&
            gridnew(nth, 0.0, thmax, 0, 1.0, tcn, tmn) && gridnew(nz, 0.0, p.alx3, 0, 1.0, zcn, zmn) &&
            gridnew(nr, p.rint, p.rext, p.istr, p.str, rcn, rmn);
  if (!ok) return set_error("tc_regrid_field: interp.F90's gridnew builds ISTR 0 and 6 only");
  for (auto *v : {&tco, &tmo}) extrapolate_ends(*v, ntho);
  for (auto *v : {&zco, &zmo}) extrapolate_ends(*v, nzo);
  for (auto *v : {&rco, &rmo}) extrapolate_ends(*v, nro);
  // interp.F90:75-112
  const bool is_qt = field == TC_QT, is_qr = field == TC_QR;
  const std::vector<Bracket> bt = brackets(is_qt ? tco : tmo, ntho, is_qt ? tcn : tmn, c->nthm);
  const std::vector<Bracket> bz = brackets(is_qr ? zco : zmo, nzo, is_qr ? zcn : zmn, c->nzm);
  const std::vector<Bracket> br =
      brackets(field == TC_QZ ? rco : rmo, nro, field == TC_QZ ? rcn : rmn, nr);
  // the old buffer with its extrapolated halos, interp.F90:113-132
  const int ak = nro + 2, aj = nzo + 1;
  std::vector<double> a((size_t)ak * aj * (size_t)(ntho + 1), 0.0);
  auto A = [&](int k, int j, int i) -> double & {
    return a[(size_t)k + (size_t)ak * ((size_t)j + (size_t)aj * (size_t)i)];
  };
  for (int i = 1; i <= nthom; i++)
    for (int j = 1; j <= nzom; j++)
      for (int k = 1; k <= nro; k++)
        A(k, j, i) = old[(size_t)(k - 1) + (size_t)nro * ((size_t)(j - 1) + (size_t)nzom * (i - 1))];
  for (int i = 1; i <= nthom; i++)
    for (int j = 1; j <= nzom; j++) {
      A(0, j, i) = 2.0 * A(1, j, i) - A(2, j, i);
      A(nro + 1, j, i) = 2.0 * A(nro, j, i) - A(nro - 1, j, i);
    }
  for (int j = 1; j <= nzom; j++)
    for (int k = 1; k <= nro; k++) {
      A(k, j, 0) = 2.0 * A(k, j, 1) - A(k, j, 2);
      A(k, j, nthom + 1) = 2.0 * A(k, j, nthom) - A(k, j, nthom - 1);
    }
  for (int i = 0; i <= nthom + 1; i++)
    for (int k = 1; k <= nro; k++) {
      A(k, 0, i) = 2.0 * A(k, 1, i) - A(k, 2, i);
      A(k, nzom + 1, i) = 2.0 * A(k, nzom, i) - A(k, nzom - 1, i);
    }
  for (int i = 0; i <= nthom + 1; i++)
    for (int j = 0; j <= nzom + 1; j++) {
      if (i >= 1 && i <= nthom && j >= 1 && j <= nzom) continue;
      A(0, j, i) = 2.0 * A(1, j, i) - A(2, j, i);
      A(nro + 1, j, i) = 2.0 * A(nro, j, i) - A(nro - 1, j, i);
    }
  // device: buffer + brackets, one launch, then free
  double *da = nullptr;
  Bracket *db = nullptr;
  const size_t nb = bt.size() + bz.size() + br.size();
  if (hipMalloc((void **)&da, sizeof(double) * a.size()) != hipSuccess)
    return set_error("tc_regrid_field: out of device memory");
  if (hipMalloc((void **)&db, sizeof(Bracket) * nb) != hipSuccess) {
    hipFree(da);
    return set_error("tc_regrid_field: out of device memory");
  }
  int rc = 0;
  if (hipMemcpyAsync(da, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(db, bt.data(), sizeof(Bracket) * bt.size(), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(db + bt.size(), bz.data(), sizeof(Bracket) * bz.size(), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess ||
      hipMemcpyAsync(db + bt.size() + bz.size(), br.data(), sizeof(Bracket) * br.size(),
                     hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    rc = set_error("tc_regrid_field: copy to device failed");
  } else {
    const long long n = c->P * (long long)nr;
    hipLaunchKernelGGL(k_regrid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, da, ak,
                       aj, db, db + bt.size(), db + bt.size() + bz.size(), c->nthm, c->nzm, nr,
                       c->f[field]);
    rc = check_launch();
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess)
      rc = set_error("tc_regrid_field: kernel failed");
  }
  hipFree(db);
  hipFree(da);
  if (!rc) c->cfl_valid = false;
  return rc;
}

}  // namespace tc
