// Pressure-Poisson solve of SolvePressureCorrection.F90:100-163 on one MI355X.
//
// The reference transforms z (r2c) then theta (c2c), solves one complex
// radial tridiagonal per (kz, ktheta) and transforms back, with eight pencil
// transposes in between.  Here the field never leaves its [r][z][theta]
// layout:
//   1. r2c along theta (contiguous rows; two real rows packed into one
//      complex FFT), giving spec[r][z][ktheta], ktheta = 0..nthm/2;
//   2. c2c along z on tiles of ktheta columns staged in LDS;
//   3. Thomas along r per (kz, ktheta) mode, lanes across ktheta, with the
//      1/(nthm*nzm) normalisation folded into the right-hand side;
//   4. inverse c2c along z, then c2r along theta back into dph.
// The 2-D DFT is separable, so the order of the two directions only changes
// rounding.  The FFTs are Stockham autosort passes in LDS with radix 8/4/2/3/5
// butterflies and twiddles from a host-computed table.
#include <cmath>

#include "tc_internal.h"

namespace tc {

#define DEVI __device__ __forceinline__

DEVI double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
DEVI double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
DEVI double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by (s * i), s = +-1
DEVI double2 cmul_si(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

// In-register DFT of length R with exponent sign sg (-1 forward, +1 backward).
template <int R>
DEVI void dft(double2 *v, double sg);

template <>
DEVI void dft<2>(double2 *v, double) {
  const double2 a = cadd(v[0], v[1]), b = csub(v[0], v[1]);
  v[0] = a;
  v[1] = b;
}

template <>
DEVI void dft<3>(double2 *v, double sg) {
  const double s60 = 0.86602540378443864676;
  const double2 t1 = cadd(v[1], v[2]);
  const double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  const double2 d = csub(v[1], v[2]);
  const double2 t3 = make_double2(d.x * s60, d.y * s60);
  const double2 it3 = cmul_si(t3, sg);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, it3);
  v[2] = csub(t2, it3);
}

template <>
DEVI void dft<4>(double2 *v, double sg) {
  const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  const double2 b0 = cadd(v[1], v[3]), b1 = cmul_si(csub(v[1], v[3]), sg);
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <>
DEVI void dft<5>(double2 *v, double sg) {
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
  const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const double2 x0 = cadd(v[0], cadd(t1, t2));
  const double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  const double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  const double2 b1 = cmul_si(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sg);
  const double2 b2 = cmul_si(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sg);
  v[0] = x0;
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <>
DEVI void dft<8>(double2 *v, double sg) {
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e, sg);
  dft<4>(o, sg);
  const double c = 0.70710678118654752440;
  // o[q] *= exp(sg*2*pi*i*q/8)
  o[1] = make_double2(c * (o[1].x - sg * o[1].y), c * (o[1].y + sg * o[1].x));
  o[2] = cmul_si(o[2], sg);
  o[3] = make_double2(c * (-o[3].x - sg * o[3].y), c * (-o[3].y + sg * o[3].x));
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[q] = cadd(e[q], o[q]);
    v[q + 4] = csub(e[q], o[q]);
  }
}

// One Stockham pass of radix R over a line of n points in LDS, element e at
// buf[e*stride].  In place: every thread reads the inputs of its butterflies
// into registers, barrier, writes, barrier.  Ns is the product of the radices
// of the earlier passes.  ME bounds the points a thread holds per pass
// (n <= ME*T), so the register arrays have a compile-time size.  All threads
// of the block must call.
template <int R, int ME>
DEVI void fft_pass(double2 *buf, int stride, int n, int Ns, int t, int T,
                   const double2 *__restrict__ tw, double sg) {
  constexpr int MB = (ME + R - 1) / R;
  const int nb = n / R;
  const int twstep = n / (Ns * R);
  double2 v[MB][R];
#pragma unroll
  for (int q = 0; q < MB; q++) {
    const int b = t + q * T;
    if (b < nb) {
      const int k = b % Ns;
#pragma unroll
      for (int r = 0; r < R; r++) {
        double2 x = buf[(b + r * nb) * stride];
        if (r > 0 && k > 0) {
          double2 w = tw[r * k * twstep];
          w.y = (sg < 0) ? w.y : -w.y;
          x = cmul(x, w);
        }
        v[q][r] = x;
      }
      dft<R>(v[q], sg);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; q++) {
    const int b = t + q * T;
    if (b < nb) {
      const int k = b % Ns;
      const int base = (b / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[(base + r * Ns) * stride] = v[q][r];
    }
  }
  __syncthreads();
}

template <int ME>
DEVI void fft_line(double2 *buf, int stride, const FFTPlan &pl, int t, int T,
                   const double2 *__restrict__ tw, double sg) {
  int Ns = 1;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.radix[s];
    switch (R) {
      case 8: fft_pass<8, ME>(buf, stride, pl.n, Ns, t, T, tw, sg); break;
      case 4: fft_pass<4, ME>(buf, stride, pl.n, Ns, t, T, tw, sg); break;
      case 2: fft_pass<2, ME>(buf, stride, pl.n, Ns, t, T, tw, sg); break;
      case 3: fft_pass<3, ME>(buf, stride, pl.n, Ns, t, T, tw, sg); break;
      default: fft_pass<5, ME>(buf, stride, pl.n, Ns, t, T, tw, sg); break;
    }
    Ns *= R;
  }
}

// r2c along theta for row pairs: rows 2p and 2p+1 of the real input become
// the real and imaginary parts of one complex FFT.  Row r of the real input
// is at in + r*nthm, row r of the half spectrum at spec + r*nth2.
template <int ME>
__global__ __launch_bounds__(256) void k_fft_theta_fwd(const double *__restrict__ in,
                                                       double2 *__restrict__ spec, FFTPlan pl,
                                                       const double2 *__restrict__ tw,
                                                       long long nrows, int nth2, int sstr) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = 256 / L;
  const int ln = threadIdx.x / T, t = threadIdx.x - ln * T;
  const long long pair = (long long)blockIdx.x * L + ln;
  const long long r0 = 2 * pair, r1 = r0 + 1;
  const bool v0 = r0 < nrows, v1 = r1 < nrows;
  double2 *buf = lds + ln * n;
  const double *a = in + r0 * n, *b = in + r1 * n;
  for (int e = t; e < n; e += T) buf[e] = make_double2(v0 ? a[e] : 0.0, v1 ? b[e] : 0.0);
  __syncthreads();
  fft_line<ME>(buf, 1, pl, t, T, tw, -1.0);
  for (int k = t; k < nth2; k += T) {
    const double2 z = buf[k];
    const double2 zc = buf[(n - k) % n];
    // A = (Z + conj(Zc))/2, B = (Z - conj(Zc))/(2i)
    const double2 A = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
    const double2 B = make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
    if (v0) spec[r0 * sstr + k] = A;
    if (v1) spec[r1 * sstr + k] = B;
  }
}

// c2r along theta for row pairs; the imaginary parts of ktheta = 0 and
// (n even) ktheta = n/2 are ignored, as a c2r transform does.
template <int ME>
__global__ __launch_bounds__(256) void k_fft_theta_inv(const double2 *__restrict__ spec,
                                                       double *__restrict__ out, FFTPlan pl,
                                                       const double2 *__restrict__ tw,
                                                       long long nrows, int nth2, int sstr) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = 256 / L;
  const int ln = threadIdx.x / T, t = threadIdx.x - ln * T;
  const long long pair = (long long)blockIdx.x * L + ln;
  const long long r0 = 2 * pair, r1 = r0 + 1;
  const bool v0 = r0 < nrows, v1 = r1 < nrows;
  double2 *buf = lds + ln * n;
  for (int k = t; k < n; k += T) {
    const bool lo = (k < nth2);
    const int kk = lo ? k : n - k;
    double2 A = v0 ? spec[r0 * sstr + kk] : make_double2(0.0, 0.0);
    double2 B = v1 ? spec[r1 * sstr + kk] : make_double2(0.0, 0.0);
    if (!lo) {
      A.y = -A.y;
      B.y = -B.y;
    }
    if (k == 0 || 2 * k == n) {
      A.y = 0.0;
      B.y = 0.0;
    }
    buf[k] = make_double2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  fft_line<ME>(buf, 1, pl, t, T, tw, 1.0);
  for (int e = t; e < n; e += T) {
    const double2 z = buf[e];
    if (v0) out[r0 * n + e] = z.x;
    if (v1) out[r1 * n + e] = z.y;
  }
}

// Workgroups are dispatched round-robin over the XCDs; hand each XCD a
// contiguous run of tiles so tiles sharing cache lines share one L2.
DEVI int xcd_tile(int b, int nb) {
  constexpr int NX = 8;
  const int q = nb / NX, r = nb - q * NX;
  const int x = b % NX, local = b / NX;
  return x * q + (x < r ? x : r) + local;
}

// c2c along z for a tile of pl.lines ktheta columns of one r-plane.  The LDS
// tile keeps the global row-major order [z][column]; lane = column fastest,
// so global and LDS accesses are both contiguous across lanes.  Spectral rows
// are sstr complex apart (a multiple of 8, i.e. of 128 bytes).
template <int ME, int NT>
__global__ __launch_bounds__(NT) void k_fft_z(double2 *__restrict__ spec, FFTPlan pl,
                                              const double2 *__restrict__ tw, int nth2, int sstr,
                                              int ntile, double sg) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = NT / L;
  const int c = threadIdx.x % L, t = threadIdx.x / L;
  const int id = xcd_tile(blockIdx.x, gridDim.x);
  const int c0 = (id % ntile) * L;
  double2 *plane = spec + (long long)(id / ntile) * n * sstr;
  for (int e = threadIdx.x; e < L * n; e += NT) {
    const int z = e / L, cc = e - z * L;
    lds[e] = (c0 + cc < nth2) ? plane[(long long)z * sstr + c0 + cc] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_line<ME>(lds + c, L, pl, t, T, tw, sg);
  for (int e = threadIdx.x; e < L * n; e += NT) {
    const int z = e / L, cc = e - z * L;
    if (c0 + cc < nth2) plane[(long long)z * sstr + c0 + cc] = lds[e];
  }
}

// SolvePressureCorrection.F90:127-148: radial tridiagonal per (kz, ktheta)
// mode, sub-diagonal amph, super-diagonal apph and diagonal
// acph - ak3(kz) - ak1(ktheta)*usrm^2.  Thomas without pivoting (the rows are
// diagonally dominant).  The singular (0,0) Neumann mode is solved on rows
// 1..nrm-1 with phi(nrm) = 0.
__global__ __launch_bounds__(256) void k_poisson_tridiag(GridDev g, double2 *__restrict__ spec,
                                                         double *__restrict__ cpw,
                                                         double inv_norm) {
  constexpr int CH = 8;  // rows loaded ahead of the recurrence
  const long long M = (long long)g.nzm * g.sstr;
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int kz = (int)(m / g.sstr), kt = (int)(m - (long long)kz * g.sstr);
  if (kt >= g.nth2) return;  // row padding
  const double sz = g.ak3[kz + 1], st = g.ak1[kt + 1];
  const int nn = (m == 0) ? g.nrm - 1 : g.nrm;
  // forward elimination
  double b = g.acph[1] - sz - st * (g.usrm[1] * g.usrm[1]);
  double rden = 1.0 / b;
  double cp = g.apph[1] * rden;
  double2 d = spec[m];
  double2 dp = make_double2(d.x * inv_norm * rden, d.y * inv_norm * rden);
  cpw[m] = cp;
  spec[m] = dp;
  for (int k0 = 2; k0 <= nn; k0 += CH) {
    double2 dv[CH];
#pragma unroll
    for (int u = 0; u < CH; u++)
      if (k0 + u <= nn) dv[u] = spec[(long long)(k0 + u - 1) * M + m];
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int k = k0 + u;
      if (k <= nn) {
        const long long o = (long long)(k - 1) * M + m;
        const double a = g.amph[k];
        b = g.acph[k] - sz - st * (g.usrm[k] * g.usrm[k]);
        rden = 1.0 / (b - a * cp);
        dp = make_double2((dv[u].x * inv_norm - a * dp.x) * rden, (dv[u].y * inv_norm - a * dp.y) * rden);
        spec[o] = dp;
        if (k < nn) {
          cp = g.apph[k] * rden;
          cpw[o] = cp;
        }
      }
    }
  }
  // back substitution
  double2 x = dp;
  for (int k0 = nn - 1; k0 >= 1; k0 -= CH) {
    double2 yv[CH];
    double cv[CH];
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int k = k0 - u;
      if (k >= 1) {
        const long long o = (long long)(k - 1) * M + m;
        yv[u] = spec[o];
        cv[u] = cpw[o];
      }
    }
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const int k = k0 - u;
      if (k >= 1) {
        x = make_double2(yv[u].x - cv[u] * x.x, yv[u].y - cv[u] * x.y);
        spec[(long long)(k - 1) * M + m] = x;
      }
    }
  }
  if (m == 0) spec[(long long)(g.nrm - 1) * M] = make_double2(0.0, 0.0);
}

std::string make_fft_plan(int n, FFTPlan &pl, int lds_bytes, int nt) {
  pl = FFTPlan();
  pl.n = n;
  pl.nt = nt;
  int m = n;
  while (m % 8 == 0 && m > 1) { pl.radix[pl.nst++] = 8; m /= 8; }
  while (m % 4 == 0 && m > 1) { pl.radix[pl.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { pl.radix[pl.nst++] = 2; m /= 2; }
  while (m % 3 == 0 && m > 1) { pl.radix[pl.nst++] = 3; m /= 3; }
  while (m % 5 == 0 && m > 1) { pl.radix[pl.nst++] = 5; m /= 5; }
  if (m != 1) return "FFT length " + std::to_string(n) + " has a prime factor other than 2, 3, 5";
  // lines per 256-thread block: as many as the LDS budget allows (at most
  // 64); then the points per thread per pass, ME, from {4, 8, 16, 32}
  int L = 64;
  while (L > 1 && (long long)L * n * 16 > lds_bytes) L /= 2;
  while (L > 1 && nt / L < 2) L /= 2;
  for (;;) {
    const int T = nt / L;
    const int need = (n + T - 1) / T;
    int me = 4;
    while (me < need) me *= 2;
    if (me <= 32) {
      pl.me = me;
      break;
    }
    if (L == 1) return "FFT length " + std::to_string(n) + " is too long for one block";
    L /= 2;
  }
  if ((long long)L * n * 16 > lds_bytes)
    return "FFT length " + std::to_string(n) + " does not fit the LDS budget";
  pl.lines = L;
  return std::string();
}

void fill_twiddles(int n, std::vector<double> &tw) {
  tw.resize(2 * (size_t)n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; m++) {
    const long double a = twopi * (long double)m / (long double)n;
    tw[2 * (size_t)m] = (double)cosl(a);
    tw[2 * (size_t)m + 1] = (double)(-sinl(a));
  }
}

template <int ME>
static void launch_theta(tc_ctx *c, bool fwd, long long nrows, unsigned blocks, size_t lds) {
  if (fwd)
    hipLaunchKernelGGL(k_fft_theta_fwd<ME>, dim3(blocks), dim3(256), lds, c->stream, c->f[TC_DPH],
                       c->spec, c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
  else
    hipLaunchKernelGGL(k_fft_theta_inv<ME>, dim3(blocks), dim3(256), lds, c->stream, c->spec,
                       c->f[TC_DPH], c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
}

static void theta_pass(tc_ctx *c, bool fwd) {
  const long long nrows = (long long)c->nrm * c->nzm;
  const long long npairs = (nrows + 1) / 2;
  const FFTPlan &pt = c->plan_t;
  const size_t lds = (size_t)pt.lines * pt.n * sizeof(double2);
  const unsigned blocks = (unsigned)((npairs + pt.lines - 1) / pt.lines);
  switch (pt.me) {
    case 4: launch_theta<4>(c, fwd, nrows, blocks, lds); break;
    case 8: launch_theta<8>(c, fwd, nrows, blocks, lds); break;
    case 16: launch_theta<16>(c, fwd, nrows, blocks, lds); break;
    default: launch_theta<32>(c, fwd, nrows, blocks, lds); break;
  }
}

template <int NT>
static void z_pass_nt(tc_ctx *c, double sg) {
  const FFTPlan &pz = c->plan_z;
  const size_t lds = (size_t)pz.lines * pz.n * sizeof(double2);
  const int ntile = (c->nth2 + pz.lines - 1) / pz.lines;
  const dim3 grid((unsigned)(ntile * c->nrm));
#define TC_ZLAUNCH(ME)                                                                          \
  hipLaunchKernelGGL((k_fft_z<ME, NT>), grid, dim3(NT), lds, c->stream, c->spec, pz, c->tw_z, \
                     c->nth2, c->sstr, ntile, sg)
  switch (pz.me) {
    case 4: TC_ZLAUNCH(4); break;
    case 8: TC_ZLAUNCH(8); break;
    case 16: TC_ZLAUNCH(16); break;
    default: TC_ZLAUNCH(32); break;
  }
#undef TC_ZLAUNCH
}

static void z_pass(tc_ctx *c, double sg) {
  if (c->plan_z.nt == 512)
    z_pass_nt<512>(c, sg);
  else
    z_pass_nt<256>(c, sg);
}

void launch_poisson(tc_ctx *c) {
  const long long M = (long long)c->nzm * c->sstr;
  timer_begin(c, "fft_theta_fwd");
  theta_pass(c, true);
  timer_end(c);
  timer_begin(c, "fft_z_fwd");
  z_pass(c, -1.0);
  timer_end(c);
  timer_begin(c, "poisson_tridiag");
  hipLaunchKernelGGL(k_poisson_tridiag, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                     c->gd, c->spec, c->cpw, 1.0 / ((double)c->nthm * (double)c->nzm));
  timer_end(c);
  timer_begin(c, "fft_z_inv");
  z_pass(c, 1.0);
  timer_end(c);
  timer_begin(c, "fft_theta_inv");
  theta_pass(c, false);
  timer_end(c);
}

}  // namespace tc
