// MI355X (gfx950) kernels of the statistics pass: CalcStats with the
// vorticity of CalcVorticity (StatRoutines.F90:12-71, CalcVorticity.F90:11-88)
// and CalcDissipation (CalcDissipation.F90:11-153).
//
// One 256-thread block owns one (r-plane k, z-row j) theta line: its
// threads stride along theta, each sums its points, a fixed butterfly
// reduces each wave, the four wave sums are added in order, and the block
// adds the result to the (k, j) accumulators it alone owns.  So the
// accumulators need no atomics and the sums are the same run to run.  The
// vorticity is evaluated at the point and never stored.
//
// Accumulators: TC_NSTATS arrays of (nrm, nzm) in Fortran order (k fastest),
// in WriteStats order (StatRoutines.F90:97-152), so a download is the
// reference's stat_arrays as they are.
#include "tc_internal.h"

namespace tc {

#define DEVI __device__ __forceinline__

namespace {

struct Fields {
  const double *__restrict__ qt, *__restrict__ qr, *__restrict__ qz, *__restrict__ pr;
};

constexpr int NS = TC_NSTATS;  // accumulators
constexpr int NV = NS + 1;     // + the pvol-weighted dissipation of nute

// One 256-thread block per (k, j) theta line: lanes stride along theta by
// 256, each wave reduces with a fixed butterfly, and the four wave sums are
// added in wave order.
template <bool STATS, bool DISS>
__global__ __launch_bounds__(256) void k_stats(GridDev g, Fields f, double *__restrict__ acc,
                                               double *__restrict__ part, double ren) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x;
  const int k = blockIdx.y + 1;
  const int nthm = g.nthm, nzm = g.nzm, nrm = g.nrm;
  const long long P = g.P;
  const int kp = k + 1;                       // qr face above the cell
  const int kpp = k < nrm ? k + 1 : nrm;      // kpv(k), CreateGrid.F90:39-45
  const int kmm = k > 1 ? k - 1 : 1;          // kmv(k)
  const int jp = j + 1 == nzm ? 0 : j + 1, jm = j == 0 ? nzm - 1 : j - 1;
  const double dxt = g.dxt, dxz = g.dxz, pi = M_PI;
  const double usrc_k = g.usrc[k], usrc_kp = g.usrc[kp], usrc_kpp = g.usrc[kpp];
  const double usrc_kmm = g.usrc[kmm];
  const double usrm_k = g.usrm[k], usrm_kpp = g.usrm[kpp], usrm_kmm = g.usrm[kmm];
  const double rm_k = g.rm[k], rm_kpp = g.rm[kpp], rm_kmm = g.rm[kmm];
  const double udx2mm = g.udx2mm[k], udx2cc = g.udx2cc[k], pvol = g.pvol[k];
  const double usnthm = 1.0 / nthm;
  auto at = [&](const double *a, int kk, int jj, int ii) {
    return a[(long long)(kk - 1) * P + (long long)jj * nthm + ii];
  };

  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) s[q] = 0.0;

  for (int i = threadIdx.x; i < nthm; i += 256) {
    const int ip = i + 1 == nthm ? 0 : i + 1, im = i == 0 ? nthm - 1 : i - 1;
    if (STATS) {
      // CalcVorticity.F90:41-79
      const double durdz = ((at(f.qr, kp, jp, i) * usrc_kp + at(f.qr, k, jp, i) * usrc_k) -
                            (at(f.qr, kp, jm, i) * usrc_kp + at(f.qr, k, jm, i) * usrc_k)) * dxz * 0.25;
      const double duzdr = ((at(f.qz, kpp, j, i) + at(f.qz, kpp, jp, i)) -
                            (at(f.qz, kmm, j, i) + at(f.qz, kmm, jp, i))) * 0.5 * udx2mm;
      const double vt = durdz - duzdr;
      const double duzdt = ((at(f.qz, k, j, ip) + at(f.qz, k, jp, ip)) -
                            (at(f.qz, k, j, im) + at(f.qz, k, jp, im))) * 0.25 * dxt * usrm_k;
      const double dutdz = ((at(f.qt, k, jp, i) + at(f.qt, k, jp, ip)) -
                            (at(f.qt, k, jm, i) + at(f.qt, k, jm, ip))) * 0.25 * dxz;
      const double vr = duzdt - dutdz;
      const double dutdr = ((at(f.qt, kpp, j, i) + at(f.qt, kpp, j, ip)) * rm_kpp -
                            (at(f.qt, kmm, j, i) + at(f.qt, kmm, j, ip)) * rm_kmm) * 0.5 * udx2mm * usrm_k;
      const double durdt = ((at(f.qr, k, j, ip) * usrc_k + at(f.qr, kp, j, ip) * usrc_kp) -
                            (at(f.qr, k, j, im) * usrc_k + at(f.qr, kp, j, im) * usrc_kp)) * dxt * 0.25;
      const double vz = dutdr - durdt;
      // StatRoutines.F90:33-63 (kp there is kpv(k))
      const double qt = at(f.qt, k, j, i), qr = at(f.qr, k, j, i), qz = at(f.qz, k, j, i);
      const double pr = at(f.pr, k, j, i);
      const double qtip = at(f.qt, k, j, ip), qzjp = at(f.qz, k, jp, i);
      const double qrjp = at(f.qr, k, jp, i), qrkv = at(f.qr, kpp, j, i);
      s[TC_ST_VT_MEAN] += qt * usnthm;
      s[TC_ST_VR_MEAN] += qr * usnthm * usrc_k;
      s[TC_ST_VZ_MEAN] += qz * usnthm;
      s[TC_ST_WT_MEAN] += vt * usnthm;
      s[TC_ST_WR_MEAN] += vr * usnthm;
      s[TC_ST_WZ_MEAN] += vz * usnthm;
      s[TC_ST_VT_RMS] += qt * qt * usnthm;
      s[TC_ST_VR_RMS] += (qr * usrc_k) * (qr * usrc_k) * usnthm;
      s[TC_ST_VZ_RMS] += qz * qz * usnthm;
      s[TC_ST_WT_RMS] += vt * vt * usnthm;
      s[TC_ST_WR_RMS] += vr * vr * usnthm;
      s[TC_ST_WZ_RMS] += vz * vz * usnthm;
      s[TC_ST_PR_MEAN] += pr * usnthm;
      s[TC_ST_PR_RMS] += pr * pr * usnthm;
      const double vrc = qr * usrc_k + qrkv * usrc_kpp;
      s[TC_ST_VTVR_MEAN] += (qt + qtip) * vrc * usnthm * 0.25;
      // the reference pairs qz(k,j,i) with qr(k,j+1,i) here
      s[TC_ST_VTVZ_MEAN] += (qt + qtip) * (qz + qrjp) * usnthm * 0.25;
      s[TC_ST_VRVZ_MEAN] += (qz + qzjp) * vrc * usnthm * 0.25;
    }
    if (DISS) {
      // CalcDissipation.F90:62-133
      const double e11 = (at(f.qt, k, j, ip) - at(f.qt, k, j, im)) * dxt * usrm_k * 0.5 +
                         at(f.qr, k, j, i) * (usrc_k * usrc_k);
      const double e12 = (at(f.qt, kpp, j, i) * usrm_kpp - at(f.qt, kmm, j, i) * usrm_kmm) * udx2mm * rm_k +
                         (at(f.qr, k, j, ip) - at(f.qr, k, j, im)) * usrc_k * dxt * 0.5;
      const double e13 = (at(f.qt, k, jp, i) - at(f.qt, k, jm, i)) * 0.5 * dxz +
                         (at(f.qz, k, j, ip) - at(f.qz, k, j, im)) * 0.5 * dxt * usrm_k;
      const double e22 = (at(f.qr, kp, j, i) * usrc_kp - at(f.qr, kmm, j, i) * usrc_kmm) * udx2cc;
      const double e23 = (at(f.qz, kpp, j, i) - at(f.qz, kmm, j, i)) * udx2mm +
                         (at(f.qr, k, jp, i) - at(f.qr, k, jm, i)) * usrc_k * dxz * 0.5;
      const double e33 = (at(f.qz, k, jp, i) - at(f.qz, k, jm, i)) * dxz * 0.5;
      const double d = 2.0 * (e11 * e11 + e22 * e22 + e33 * e33) + e12 * e12 + e13 * e13 + e23 * e23;
      s[TC_ST_DISSTE] += d / ren / (double)nthm;
      s[NS] += d * pvol * pi;
    }
  }
  // fixed butterfly over each wave, then the four waves in order
  __shared__ double sh[4][NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double v = s[q];
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) v += __shfl_xor(v, w, 64);
    if (lane == 0) sh[wave][q] = v;
  }
  __syncthreads();
  // thread q adds accumulator q of this (k, j)
  const long long kj = (long long)(k - 1) + (long long)nrm * j;
  const long long M = (long long)nrm * nzm;
  const int q = threadIdx.x;
  if (q < NV) {
    const double v = ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q];
    if (q < NS) {
      const bool mine = (q == TC_ST_DISSTE) ? DISS : STATS;
      if (mine) acc[q * M + kj] += v;
    } else if (DISS) {
      part[kj] = v;
    }
  }
}

// Sum of the n per-line dissipation partials in a fixed order (one block).
__global__ __launch_bounds__(256) void k_sum_fixed(const double *__restrict__ part, long long n,
                                                   double *__restrict__ out) {
  __shared__ double sh[256];
  double v = 0.0;
  for (long long b = threadIdx.x; b < n; b += 256) v += part[b];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// DumpRCuts.F90:28-36 at r-plane k: vt, vr (mean of the two faces), vz;
// out[q] is (nzm, nthm) in Fortran order (z fastest).
__global__ __launch_bounds__(256) void k_rcut(GridDev g, Fields f, int k, double *__restrict__ out) {
  const long long l = (long long)blockIdx.x * 256 + threadIdx.x;
  if (l >= g.P) return;
  const int j = (int)(l / g.nthm), i = (int)(l - (long long)j * g.nthm);
  const long long o = (long long)(k - 1) * g.P + l;
  const long long t = (long long)i * g.nzm + j;
  out[t] = f.qt[o];
  out[g.P + t] = (f.qr[o] * g.usrc[k] + f.qr[o + g.P] * g.usrc[k + 1]) * 0.5;
  out[2 * g.P + t] = f.qz[o];
}

// DumpThCuts.F90:33-44: the theta = 1 cut; out[q] is (nrm, nzm), r fastest.
__global__ __launch_bounds__(256) void k_thcut(GridDev g, Fields f, double *__restrict__ out) {
  const long long m = (long long)g.nrm * g.nzm;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= m) return;
  const int j = (int)(t / g.nrm), k = (int)(t - (long long)j * g.nrm) + 1;
  const long long o = (long long)(k - 1) * g.P + (long long)j * g.nthm;
  out[t] = f.qt[o];
  out[m + t] = f.qr[o] * g.usrc[k];
  out[2 * m + t] = f.qz[o];
}

// DumpStreamwiseAvg.F90:33-44: theta-means of vt, vr, vz, one wave per
// (k, j) line as in k_stats; out[q] is (nrm, nzm), r fastest.
__global__ __launch_bounds__(256) void k_theta_avg(GridDev g, Fields f, double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.y + 1;
  if (j >= g.nzm) return;
  const double usnthm = 1.0 / (double)g.nthm, usrc = g.usrc[k];
  const long long row = (long long)(k - 1) * g.P + (long long)j * g.nthm;
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = lane; i < g.nthm; i += 64) {
    s[0] += f.qt[row + i] * usnthm;
    s[1] += f.qr[row + i] * usrc * usnthm;
    s[2] += f.qz[row + i] * usnthm;
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {
    double v = s[q];
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) v += __shfl_xor(v, w, 64);
    s[q] = v;
  }
  const long long m = (long long)g.nrm * g.nzm;
  if (lane < 3) out[lane * m + (long long)(k - 1) + (long long)g.nrm * j] = s[lane];
}

}  // namespace

// Slab extraction into c->dump (3 arrays); kind 0: r-cut at plane k,
// 1: theta cut, 2: theta-means.
void launch_dump(tc_ctx *c, int kind, int k) {
  Fields f{c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_PR]};
  const long long m = (long long)c->nrm * c->nzm;
  if (kind == 0)
    hipLaunchKernelGGL(k_rcut, dim3((unsigned)((c->P + 255) / 256)), dim3(256), 0, c->stream, c->gd,
                       f, k, c->dump);
  else if (kind == 1)
    hipLaunchKernelGGL(k_thcut, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream, c->gd, f,
                       c->dump);
  else
    hipLaunchKernelGGL(k_theta_avg, dim3((unsigned)((c->nzm + 3) / 4), (unsigned)c->nrm), dim3(256),
                       0, c->stream, c->gd, f, c->dump);
}

// Statistics pass over the current fields.  stats: CalcStats (with the
// vorticity); diss: CalcDissipation, whose volume sum lands in c->red_out[0].
void launch_stats(tc_ctx *c, bool stats, bool diss) {
  timer_begin(c, "stats");
  const dim3 grid((unsigned)c->nzm, (unsigned)c->nrm);
  Fields f{c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], c->f[TC_PR]};
  const double ren = c->p.ren;
  if (stats && diss)
    hipLaunchKernelGGL((k_stats<true, true>), grid, dim3(256), 0, c->stream, c->gd, f, c->stats,
                       c->stat_part, ren);
  else if (stats)
    hipLaunchKernelGGL((k_stats<true, false>), grid, dim3(256), 0, c->stream, c->gd, f, c->stats,
                       c->stat_part, ren);
  else
    hipLaunchKernelGGL((k_stats<false, true>), grid, dim3(256), 0, c->stream, c->gd, f, c->stats,
                       c->stat_part, ren);
  if (diss) {
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, c->stream, c->stat_part,
                       (long long)c->nrm * c->nzm, c->red_out);
    hipMemcpyAsync(c->red_host, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  }
  timer_end(c);
}

}  // namespace tc
