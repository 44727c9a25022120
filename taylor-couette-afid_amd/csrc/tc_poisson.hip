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
#include <cstdlib>

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

// Non-temporal access for the spectrum and dph where a kernel streams each
// element exactly once (a streaming copy runs 5.8% faster with both).
DEVI double2 ldc_nt(const double2 *p) {
  return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
DEVI void stc_nt(double2 *p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

// LDS slot of element x with one padding slot after every 2^sh elements
DEVI int pidx(int x, int sh) { return x + (x >> sh); }

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

// One Stockham pass of radix R over a line of n points.  Ns is the product
// of the radices of the earlier passes; butterfly b reads points b + r*n/R
// and writes points (b/Ns)*Ns*R + b%Ns + r*Ns.  Points live in LDS, except
// that the first pass may read them from memory
// through ld(e) (FROM_G) and the last pass may write them to memory through
// st(e, value) (TO_G), which saves one LDS round trip and one barrier at
// each end of the transform.  Point e of the line is at LDS slot
// pidx(off + e*stride): one padding slot per 2^sh slots spreads the strided
// Stockham writes over the LDS banks.  ME bounds the points a thread holds per pass
// (n <= ME*T), so the register arrays have a compile-time size.  All threads
// of the block must call.
// BATCH: all loads first and unconditionally (a thread past the last
// butterfly repeats the last one and does not store it), so they are issued
// as one batch, at the cost of registers; tw[0] = 1 exactly, so k = 0 needs
// no branch.  Otherwise each butterfly loads and multiplies in turn.
template <int R, int ME, bool FROM_G, bool TO_G, bool BATCH, class Load, class Store>
DEVI void fft_pass(double2 *buf, int off, int stride, int sh, int n, int Ns, int t, int T,
                   const double2 *__restrict__ tw, double sg, Load &ld, Store &st) {
  constexpr int MB = (ME + R - 1) / R;
  const int nb = n / R;
  const int twstep = n / (Ns * R);
  double2 v[MB][R];
  if constexpr (!BATCH) {
#pragma unroll
    for (int q = 0; q < MB; q++) {
      const int b = t + q * T;
      if (b < nb) {
        const int k = b % Ns;
#pragma unroll
        for (int r = 0; r < R; r++) {
          double2 x;
          if constexpr (FROM_G) x = ld(b + r * nb);
          else x = buf[pidx(off + (b + r * nb) * stride, sh)];
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
  } else {
#pragma unroll
  for (int q = 0; q < MB; q++) {
    const int b = min(t + q * T, nb - 1);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if constexpr (FROM_G) v[q][r] = ld(b + r * nb);
      else v[q][r] = buf[pidx(off + (b + r * nb) * stride, sh)];
    }
  }
  double2 w[MB][R];
#pragma unroll
  for (int q = 0; q < MB; q++) {
    const int b = min(t + q * T, nb - 1);
    const int k = b % Ns;
#pragma unroll
    for (int r = 1; r < R; r++) w[q][r] = tw[r * k * twstep];
  }
#pragma unroll
  for (int q = 0; q < MB; q++) {
#pragma unroll
    for (int r = 1; r < R; r++) {
      double2 x = w[q][r];
      x.y = (sg < 0) ? x.y : -x.y;
      v[q][r] = cmul(v[q][r], x);
    }
    dft<R>(v[q], sg);
  }
  }
  if constexpr (!FROM_G) __syncthreads();
#pragma unroll
  for (int q = 0; q < MB; q++) {
    const int b = t + q * T;
    if (b < nb) {
      const int k = b % Ns;
      const int base = (b / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (TO_G) st(base + r * Ns, v[q][r]);
        else buf[pidx(off + (base + r * Ns) * stride, sh)] = v[q][r];
      }
    }
  }
  if constexpr (!TO_G) __syncthreads();
}

template <int ME, bool FROM_G, bool TO_G, bool BATCH, class Load, class Store>
DEVI void fft_pass_r(int R, double2 *buf, int off, int stride, int sh, int n, int Ns, int t,
                     int T, const double2 *__restrict__ tw, double sg, Load &ld, Store &st) {
#define TC_FP(RR) fft_pass<RR, ME, FROM_G, TO_G, BATCH>(buf, off, stride, sh, n, Ns, t, T, tw, sg, ld, st)
  switch (R) {
    case 8: TC_FP(8); break;
    case 4: TC_FP(4); break;
    case 2: TC_FP(2); break;
    case 3: TC_FP(3); break;
    default: TC_FP(5); break;
  }
#undef TC_FP
}

// A whole line: with FROM_G the input comes from ld(e), otherwise from LDS;
// with TO_G the output goes to st(e, value), otherwise to LDS (in natural
// order, followed by a barrier).
template <int ME, bool FROM_G, bool TO_G, bool BATCH = true, class Load, class Store>
DEVI void fft_line(double2 *buf, int off, int stride, const FFTPlan &pl, int t, int T,
                   const double2 *__restrict__ tw, double sg, Load ld, Store st) {
  const int n = pl.n, ns = pl.nst, sh = pl.padsh;
  if (ns == 1) {
    fft_pass_r<ME, FROM_G, TO_G, BATCH>(pl.radix[0], buf, off, stride, sh, n, 1, t, T, tw, sg, ld, st);
    return;
  }
  fft_pass_r<ME, FROM_G, false, BATCH>(pl.radix[0], buf, off, stride, sh, n, 1, t, T, tw, sg, ld, st);
  int Ns = pl.radix[0];
  for (int s = 1; s < ns - 1; s++) {
    fft_pass_r<ME, false, false, BATCH>(pl.radix[s], buf, off, stride, sh, n, Ns, t, T, tw, sg, ld, st);
    Ns *= pl.radix[s];
  }
  fft_pass_r<ME, false, TO_G, BATCH>(pl.radix[ns - 1], buf, off, stride, sh, n, Ns, t, T, tw, sg, ld, st);
}

// r2c along theta for row pairs: rows 2p and 2p+1 of the real input become
// the real and imaginary parts of one complex FFT.  Row r of the real input
// is at in + r*nthm, row r of the half spectrum at spec + r*sstr.
// With DIV the real rows are not read from memory: the first pass computes
// them as the local divergence of CalcLocalDivergence.F90:33-37 (row
// r = (kc-1)*nzm + j), so dph is neither written nor re-read.
struct DivIn {
  const double *qt, *qr, *qz;
  double usdt;
};

template <int ME, bool DIV, bool STAGE = false>
DEVI void theta_fwd_body(const long long pb, const double *__restrict__ in, const DivIn &dv,
                         const GridDev &g, double2 *__restrict__ spec, const FFTPlan &pl,
                         const double2 *__restrict__ tw, long long nrows, int nth2, int sstr) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = 256 / L;
  const int ln = threadIdx.x / T, t = threadIdx.x - ln * T;
  const long long pair = pb * L + ln;
  const long long r0 = 2 * pair, r1 = r0 + 1;
  const bool v0 = r0 < nrows, v1 = r1 < nrows;
  double2 *buf = lds;
  const int off = ln * n, sh = pl.padsh;
  if constexpr (DIV) {
    // per row: qt, qr(kc), qr(kc+1) = qr(kc) + P, qz(j), qz(j+1) and the metric
    struct Row {
      const double *qt, *qr, *qz, *qzp;  // qt, qr(kc), qz(j), qz(j+1) rows
      double a, b;                       // usrm(kc), udx2m(kc)
    };
    auto row = [&](long long r) {
      Row w;
      const long long rr = r < nrows ? r : 0;
      const int i = (int)(rr / g.nzm), j = (int)(rr - (long long)i * g.nzm);
      const int jp = (j + 1 == g.nzm) ? 0 : j + 1;
      w.qt = dv.qt + rr * n;
      w.qr = dv.qr + rr * n;
      w.qz = dv.qz + rr * n;
      w.qzp = dv.qz + ((long long)i * g.nzm + jp) * n;
      w.a = g.usrm[i + 1];
      w.b = g.udx2m[i + 1];
      return w;
    };
    const Row w0 = row(r0), w1 = row(r1);
    const long long P = g.P;
    auto div = [&](const Row &w, int e) {
      const int ep = (e + 1 == n) ? 0 : e + 1;
      return (((w.qt[ep] - w.qt[e]) * g.dxt + (w.qr[e + P] - w.qr[e]) * w.b) * w.a +
              (w.qzp[e] - w.qz[e]) * g.dxz) *
             dv.usdt;
    };
    // rows past nrows read row 0 (w clamped) and are zeroed after, so the
    // loads carry no condition
    auto ld = [&](int e) {
      const double d0 = div(w0, e), d1 = div(w1, e);
      return make_double2(v0 ? d0 : 0.0, v1 ? d1 : 0.0);
    };
    auto st = [&](int, double2) {};
    if constexpr (STAGE) {
      // n <= ME * T: ME rounds, the last ones clamped to element n-1
#pragma unroll
      for (int q = 0; q < ME; q++) {
        const int e = t + q * T;
        const double2 v = ld(e < n ? e : n - 1);
        if (e < n) buf[pidx(off + e, sh)] = v;
      }
      __syncthreads();
      fft_line<ME, false, false>(buf, off, 1, pl, t, T, tw, -1.0, ld, st);
    } else {
      fft_line<ME, true, false>(buf, off, 1, pl, t, T, tw, -1.0, ld, st);
    }
  } else {
    const double *a = in + r0 * n, *b = in + r1 * n;
    auto ld = [&](int e) { return make_double2(v0 ? a[e] : 0.0, v1 ? b[e] : 0.0); };
    auto st = [&](int, double2) {};
    fft_line<ME, true, false>(buf, off, 1, pl, t, T, tw, -1.0, ld, st);
  }
  for (int k = t; k < nth2; k += T) {
    const double2 z = buf[pidx(off + k, sh)];
    const double2 zc = buf[pidx(off + (n - k) % n, sh)];
    // A = (Z + conj(Zc))/2, B = (Z - conj(Zc))/(2i)
    const double2 A = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
    const double2 B = make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
    if (v0) stc_nt(spec + r0 * sstr + k, A);
    if (v1) stc_nt(spec + r1 * sstr + k, B);
  }
}

template <int ME>
__global__ __launch_bounds__(256) void k_fft_theta_fwd(const double *__restrict__ in,
                                                       double2 *__restrict__ spec, FFTPlan pl,
                                                       const double2 *__restrict__ tw,
                                                       long long nrows, int nth2, int sstr) {
  const DivIn dv{};
  const GridDev g{};
  theta_fwd_body<ME, false>(blockIdx.x, in, dv, g, spec, pl, tw, nrows, nth2, sstr);
}

template <int ME, bool STAGE>
__global__ __launch_bounds__(256) void k_div_fft_theta_fwd(DivIn dv, GridDev g,
                                                           double2 *__restrict__ spec, FFTPlan pl,
                                                           const double2 *__restrict__ tw,
                                                           long long nrows, int nth2, int sstr) {
  theta_fwd_body<ME, true, STAGE>(blockIdx.x, nullptr, dv, g, spec, pl, tw, nrows, nth2, sstr);
}

// Grouped: blockIdx.y = instance.
template <int ME>
__global__ __launch_bounds__(256) void k_div_fft_theta_fwd_g(const StepDesc *__restrict__ tab) {
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  const DivIn dv{glob(d.f[TC_QT]), glob(d.f[TC_QR]), glob(d.f[TC_QZ]), d.usdt};
  theta_fwd_body<ME, true, true>(blockIdx.x, nullptr, dv, glob(d.g), glob(d.spec), d.pt,
                                 glob(d.twt), (long long)d.g.nrm * d.g.nzm, d.g.nth2, d.g.sstr);
}

// c2r along theta for row pairs; the imaginary parts of ktheta = 0 and
// (n even) ktheta = n/2 are ignored, as a c2r transform does.  Each half
// spectrum value is read once: Z[k] = A[k] + i B[k] and
// Z[n-k] = conj(A[k]) + i conj(B[k]) go to LDS together; the last pass
// writes the two real rows straight to memory.
template <int ME>
DEVI void theta_inv_body(const long long pb, const double2 *__restrict__ spec,
                         double *__restrict__ out, const FFTPlan &pl,
                         const double2 *__restrict__ tw, long long nrows, int nth2, int sstr) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = 256 / L;
  const int ln = threadIdx.x / T, t = threadIdx.x - ln * T;
  const long long pair = pb * L + ln;
  const long long r0 = 2 * pair, r1 = r0 + 1;
  const bool v0 = r0 < nrows, v1 = r1 < nrows;
  double2 *buf = lds;
  const int off = ln * n, sh = pl.padsh;
  // T * ME >= n, so nth2 <= T * (ME/2 + 1): the loads of all rounds are
  // issued before any is used (clamped rows and columns stay in bounds)
  constexpr int KR = ME / 2 + 1;
  double2 As[KR], Bs[KR];
  const long long ra = v0 ? r0 : 0, rb = v1 ? r1 : 0;
#pragma unroll
  for (int u = 0; u < KR; u++) {
    const int k = min(t + u * T, nth2 - 1);
    As[u] = ldc_nt(spec + ra * sstr + k);
    Bs[u] = ldc_nt(spec + rb * sstr + k);
  }
#pragma unroll
  for (int u = 0; u < KR; u++) {
    const int k = t + u * T;
    if (k < nth2) {
      double2 A = v0 ? As[u] : make_double2(0.0, 0.0);
      double2 B = v1 ? Bs[u] : make_double2(0.0, 0.0);
      if (k == 0 || 2 * k == n) {
        A.y = 0.0;
        B.y = 0.0;
      }
      buf[pidx(off + k, sh)] = make_double2(A.x - B.y, A.y + B.x);
      if (k > 0 && 2 * k != n) buf[pidx(off + n - k, sh)] = make_double2(A.x + B.y, B.x - A.y);
    }
  }
  __syncthreads();
  auto ld = [&](int) { return make_double2(0.0, 0.0); };
  auto st = [&](int e, double2 z) {
    if (v0) __builtin_nontemporal_store(z.x, out + r0 * n + e);
    if (v1) __builtin_nontemporal_store(z.y, out + r1 * n + e);
  };
  // per-butterfly passes: 104 VGPRs (4 waves per SIMD) against 165 for the
  // batched ones, 354 against 367 us at 512x512x385
  fft_line<ME, false, true, false>(buf, off, 1, pl, t, T, tw, 1.0, ld, st);
}

template <int ME>
__global__ __launch_bounds__(256) void k_fft_theta_inv(const double2 *__restrict__ spec,
                                                       double *__restrict__ out, FFTPlan pl,
                                                       const double2 *__restrict__ tw,
                                                       long long nrows, int nth2, int sstr) {
  theta_inv_body<ME>(blockIdx.x, spec, out, pl, tw, nrows, nth2, sstr);
}

// Grouped: blockIdx.y = instance; output into dph.
template <int ME>
__global__ __launch_bounds__(256) void k_fft_theta_inv_g(const StepDesc *__restrict__ tab) {
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  theta_inv_body<ME>(blockIdx.x, glob(d.spec), glob(d.f[TC_DPH]), d.pt, glob(d.twt),
                     (long long)d.g.nrm * d.g.nzm, d.g.nth2, d.g.sstr);
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
DEVI void fft_z_body(const int id, double2 *__restrict__ spec, const FFTPlan &pl,
                     const double2 *__restrict__ tw, int nth2, int sstr, int ntile, double sg) {
  extern __shared__ double2 lds[];
  const int n = pl.n, L = pl.lines, T = NT / L;
  const int c = threadIdx.x % L, t = threadIdx.x / L;
  const int c0 = (id % ntile) * L;
  double2 *col = spec + (long long)(id / ntile) * n * sstr + c0 + c;
  const bool ok = c0 + c < nth2;
  // plain accesses here: with non-temporal ones the z transform ran 7%
  // slower (365 against 341 us at 512x512x385)
  auto ld = [&](int z) { return ok ? col[(long long)z * sstr] : make_double2(0.0, 0.0); };
  auto st = [&](int z, double2 v) {
    if (ok) col[(long long)z * sstr] = v;
  };
  fft_line<ME, true, true, false>(lds, c, L, pl, t, T, tw, sg, ld, st);
}

template <int ME, int NT>
__global__ __launch_bounds__(NT) void k_fft_z(double2 *__restrict__ spec, FFTPlan pl,
                                              const double2 *__restrict__ tw, int nth2, int sstr,
                                              int ntile, double sg) {
  fft_z_body<ME, NT>(xcd_tile(blockIdx.x, gridDim.x), spec, pl, tw, nth2, sstr, ntile, sg);
}

// Grouped: the XCD remap runs over the flattened (instance, tile) index.
template <int ME, int NT>
__global__ __launch_bounds__(NT) void k_fft_z_g(const StepDesc *__restrict__ tab, int ntiles,
                                                int ntile, double sg) {
  const int lin = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int inst = lin / ntiles, id = lin - inst * ntiles;
  const StepDesc &d = tab[inst];
  if (inst_stopped(d)) return;
  fft_z_body<ME, NT>(id, glob(d.spec), d.pz, glob(d.twz), d.g.nth2, d.g.sstr, ntile, sg);
}

// SolvePressureCorrection.F90:127-148: radial tridiagonal per (kz, ktheta)
// mode, sub-diagonal amph (rows 2..nrm), super-diagonal apph (rows 1..nrm-1)
// and diagonal acph - ak3(kz) - ak1(ktheta)*usrm^2.
//
// Partition method, one read and one write of the spectrum: the nrm rows of a
// line are split over PT threads (threadIdx.y) of at most PQ rows each, held
// in registers.  Each thread eliminates its rows down and up so that every
// row couples only to the partition's first and last unknowns; those 2*PT
// boundary unknowns form a tridiagonal system with unit diagonal that one
// thread per line solves in LDS; every thread then recovers its interior
// rows.  No pivoting: the rows are diagonally dominant.
//
// The singular (0,0) Neumann mode is solved on rows 1..nrm-1 with
// phi(nrm) = 0.  Its partitions are nearly singular, so block 0 stages that
// one line in LDS and solves it with a serial Thomas sweep instead.
//   blockDim = (LB lines, PT partitions)
//   LDS = LB*2PT*4 doubles (reduced systems) + 3*nrm doubles (gauge line)
//         + 4*nrm doubles (the radial coefficients amph, apph, acph, usrm^2)
template <int PQ>  // max rows per thread
DEVI void poisson_body(const long long blk, const GridDev &g, double2 *__restrict__ spec,
                       double inv_norm) {
  // [LB][2*PT][4]: A, C, D.x, D.y, each line padded by 4 doubles so that the
  // lines' rows start 8 banks apart (poisson_lds_doubles)
  extern __shared__ double red[];
  const int LB = blockDim.x, PT = blockDim.y;
  const int lx = threadIdx.x, pp = threadIdx.y;
  const int tid = pp * LB + lx, nthr = LB * PT;
  const long long M = (long long)g.nzm * g.sstr;
  const long long m = blk * LB + lx;
  const int n = g.nrm;
  const int LS = 2 * PT * 4 + 4;                                // padded line stride
  double2 *gline = (double2 *)(red + (size_t)LB * LS);            // gauge line, n values
  double *gcp = (double *)(gline + n);                          // its multipliers
  double *cam = gcp + n, *cap = cam + n, *cac = cap + n, *cus = cac + n;  // row k at [k-1]
  const bool gblock = (blk == 0);
  const int lo = (int)(((long long)pp * n) / PT), S = (int)(((long long)(pp + 1) * n) / PT) - lo;
  double ar[PQ], cr[PQ];
  double2 d[PQ];
  // this thread's rows first, as one batch of unconditional loads (clamped
  // row, last line for idle lanes), so that their latency overlaps the
  // staging below; rows past the partition and idle lanes are zeroed after
  // the barrier
  {
    const long long mm = m < M ? m : M - 1;
#pragma unroll
    for (int i = 0; i < PQ; i++) {
      const int row = min(lo + i, n - 1);
      d[i] = ldc_nt(spec + (long long)row * M + mm);
    }
  }
  if (gblock)
    for (int k = tid; k < n; k += nthr) {
      const double2 v = spec[(long long)k * M];
      gline[k] = make_double2(v.x * inv_norm, v.y * inv_norm);
    }
  // this line's wavenumber terms and the line-independent radial
  // coefficients, staged once per block: all loads issued before any wait
  const long long mv = m < M ? m : 0;
  const int kzl = (int)(mv / g.sstr), ktl = (int)(mv - (long long)kzl * g.sstr);
  const double szl = g.ak3[kzl + 1], stl = g.ak1[min(ktl, g.nth2 - 1) + 1];
  {
    double cv[2][4];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int k = min(tid + q * nthr, n - 1);
      cv[q][0] = g.amph[k + 1];
      cv[q][1] = g.apph[k + 1];
      cv[q][2] = g.acph[k + 1];
      cv[q][3] = g.usrm[k + 1];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int k = tid + q * nthr;
      if (k < n) {
        cam[k] = cv[q][0];
        cap[k] = cv[q][1];
        cac[k] = cv[q][2];
        cus[k] = cv[q][3] * cv[q][3];
      }
    }
    for (int k = tid + 2 * nthr; k < n; k += nthr) {
      cam[k] = g.amph[k + 1];
      cap[k] = g.apph[k + 1];
      cac[k] = g.acph[k + 1];
      cus[k] = g.usrm[k + 1] * g.usrm[k + 1];
    }
  }
  const bool valid = m < M && m != 0 && ktl < g.nth2;
  const double sz = valid ? szl : 0.0, st = valid ? stl : 0.0;
  // row k (1-based) of this line: a x(k-1) + b x(k) + c x(k+1) = d
  auto coef = [&](int k, double &a, double &b, double &c) {
    a = (k > 1) ? cam[k - 1] : 0.0;
    c = (k < n) ? cap[k - 1] : 0.0;
    b = cac[k - 1] - sz - st * cus[k - 1];
    if (!valid) { a = 0.0; b = 1.0; c = 0.0; }
  };
  __syncthreads();  // coefficient tables staged
#pragma unroll
  for (int i = 0; i < PQ; i++)
    d[i] = (i < S && valid) ? make_double2(d[i].x * inv_norm, d[i].y * inv_norm)
                            : make_double2(0.0, 0.0);
  // downward: rows i >= 1 become a_i x_0 + x_i + c_i x_(i+1) = d_i
#pragma unroll
  for (int i = 0; i < PQ; i++) {
    if (i < S) {
      double a, b, c;
      coef(lo + i + 1, a, b, c);
      if (i < 2) {
        const double r = 1.0 / b;
        ar[i] = a * r;
        cr[i] = c * r;
        d[i] = make_double2(d[i].x * r, d[i].y * r);
      } else {
        const double r = 1.0 / (b - a * cr[i - 1]);
        d[i] = make_double2((d[i].x - a * d[i - 1].x) * r, (d[i].y - a * d[i - 1].y) * r);
        ar[i] = -a * ar[i - 1] * r;
        cr[i] = c * r;
      }
    }
  }
  // upward: rows 1 <= i <= S-2 become a_i x_0 + x_i + c_i x_(S-1) = d_i
#pragma unroll
  for (int i = PQ - 3; i >= 1; i--) {
    if (i <= S - 3) {
      d[i] = make_double2(d[i].x - cr[i] * d[i + 1].x, d[i].y - cr[i] * d[i + 1].y);
      ar[i] = ar[i] - cr[i] * ar[i + 1];
      cr[i] = -cr[i] * cr[i + 1];
    }
  }
  // first row: A x_(-1) + x_0 + C x_(S-1) = D
  double A0 = ar[0], C0 = cr[0];
  double2 D0 = d[0];
  if (S >= 3) {
    const double r = 1.0 / (1.0 - cr[0] * ar[1]);
    D0 = make_double2((d[0].x - cr[0] * d[1].x) * r, (d[0].y - cr[0] * d[1].y) * r);
    A0 = ar[0] * r;
    C0 = -cr[0] * cr[1] * r;
  }
  double aL = 0.0, cL = 0.0;
  double2 dL = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = 1; i < PQ; i++)
    if (i == S - 1) { aL = ar[i]; cL = cr[i]; dL = d[i]; }
  double *rl = red + (size_t)lx * LS;
  rl[(2 * pp) * 4 + 0] = A0;
  rl[(2 * pp) * 4 + 1] = C0;
  rl[(2 * pp) * 4 + 2] = D0.x;
  rl[(2 * pp) * 4 + 3] = D0.y;
  rl[(2 * pp + 1) * 4 + 0] = aL;
  rl[(2 * pp + 1) * 4 + 1] = cL;
  rl[(2 * pp + 1) * 4 + 2] = dL.x;
  rl[(2 * pp + 1) * 4 + 3] = dL.y;
  __syncthreads();
  // Unit-diagonal tridiagonal system of the 2*PT boundary unknowns, solved
  // from both ends at once: thread pp = 0 eliminates rows 0..PT-1 downward
  // (x_r + c'_r x_(r+1) = d'_r), pp = 1 rows 2PT-1..PT upward (the same
  // recurrence on the reversed rows: a''_r x_(r-1) + x_r = d''_r); the two
  // meet in a 2x2 system for x_(PT-1), x_PT, and each substitutes back
  // through its own half.  Both run one instruction stream, so they share a
  // wave's issue slots, and the serial chain is half as long.  (Row r+1 is
  // read before row r is written: the reads do not wait on the stores.)
  const int R = 2 * PT, mrow = PT;
  const bool up = (pp == 1);
  if (PT == 1) {
    // one partition (nrm < 4): its two boundary rows by one thread
    if (pp == 0) {
      const double4 t = *(const double4 *)(rl), b = *(const double4 *)(rl + 4);
      const double piv = 1.0 / (1.0 - b.x * t.y);
      const double x1 = (b.z - b.x * t.z) * piv, y1 = (b.w - b.x * t.w) * piv;
      rl[2] = t.z - t.y * x1;
      rl[3] = t.w - t.y * y1;
      rl[6] = x1;
      rl[7] = y1;
    }
  } else {
  if (pp < 2) {
    const int step = up ? -1 : 1;
    int r = up ? R - 1 : 0;
    double cprev = 0.0, dxp = 0.0, dyp = 0.0;
    double4 nx = *(const double4 *)(rl + r * 4);
    for (int q = 0; q < mrow; q++, r += step) {
      const double4 cur = nx;
      if (q + 1 < mrow) nx = *(const double4 *)(rl + (r + step) * 4);
      const double sub = up ? cur.y : cur.x, sup = up ? cur.x : cur.y;
      const double piv = 1.0 / (1.0 - sub * cprev);
      cprev = sup * piv;
      dxp = (cur.z - sub * dxp) * piv;
      dyp = (cur.w - sub * dyp) * piv;
      rl[r * 4 + (up ? 0 : 1)] = cprev;
      rl[r * 4 + 2] = dxp;
      rl[r * 4 + 3] = dyp;
    }
  }
  __syncthreads();
  double xx = 0.0, xy = 0.0;
  if (pp < 2) {
    // rows PT-1 (x + c' x_PT = d') and PT (a'' x_(PT-1) + x = d'')
    const double4 t = *(const double4 *)(rl + (mrow - 1) * 4);
    const double4 b = *(const double4 *)(rl + mrow * 4);
    const double piv = 1.0 / (1.0 - t.y * b.x);
    const double x0 = (t.z - t.y * b.z) * piv, y0 = (t.w - t.y * b.w) * piv;  // x_(PT-1)
    xx = up ? b.z - b.x * x0 : x0;                                              // x_PT
    xy = up ? b.w - b.x * y0 : y0;
  }
  __syncthreads();  // both meeting rows are read before either is overwritten
  if (pp < 2) {
    const int step = up ? 1 : -1;
    int r = up ? mrow : mrow - 1;
    rl[r * 4 + 2] = xx;
    rl[r * 4 + 3] = xy;
    for (int q = 1; q < mrow; q++) {
      r += step;
      const double4 cur = *(const double4 *)(rl + r * 4);
      const double cf = up ? cur.x : cur.y;  // a''_r or c'_r
      xx = cur.z - cf * xx;
      xy = cur.w - cf * xy;
      rl[r * 4 + 2] = xx;
      rl[r * 4 + 3] = xy;
    }
  }
  }
  if (gblock && tid == nthr - 1) {
    // gauge line: Thomas on rows 1..nrm-1 (b = acph there), phi(nrm) = 0
    const int nn = n - 1;
    double b = g.acph[1] - g.ak3[1] - g.ak1[1] * (g.usrm[1] * g.usrm[1]);
    double rden = 1.0 / b;
    double cp = g.apph[1] * rden;
    double2 dp = make_double2(gline[0].x * rden, gline[0].y * rden);
    gcp[0] = cp;
    gline[0] = dp;
    for (int k = 2; k <= nn; k++) {
      const double a = g.amph[k];
      b = g.acph[k] - g.ak3[1] - g.ak1[1] * (g.usrm[k] * g.usrm[k]);
      rden = 1.0 / (b - a * cp);
      dp = make_double2((gline[k - 1].x - a * dp.x) * rden, (gline[k - 1].y - a * dp.y) * rden);
      gline[k - 1] = dp;
      cp = g.apph[k] * rden;
      gcp[k - 1] = cp;
    }
    double2 x = dp;
    for (int k = nn - 1; k >= 1; k--) {
      x = make_double2(gline[k - 1].x - gcp[k - 1] * x.x, gline[k - 1].y - gcp[k - 1] * x.y);
      gline[k - 1] = x;
    }
    gline[n - 1] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (valid) {
    const double2 x0 = make_double2(rl[(2 * pp) * 4 + 2], rl[(2 * pp) * 4 + 3]);
    const double2 xl = make_double2(rl[(2 * pp + 1) * 4 + 2], rl[(2 * pp + 1) * 4 + 3]);
#pragma unroll
    for (int i = 0; i < PQ; i++) {
      if (i < S) {
        double2 x;
        if (i == 0) x = x0;
        else if (i == S - 1) x = xl;
        else
          x = make_double2(d[i].x - ar[i] * x0.x - cr[i] * xl.x, d[i].y - ar[i] * x0.y - cr[i] * xl.y);
        stc_nt(spec + (long long)(lo + i) * M + m, x);
      }
    }
  }
  if (gblock)
    for (int k = tid; k < n; k += nthr) spec[(long long)k * M] = gline[k];
}

template <int PQ>
__global__ __launch_bounds__(512) void k_poisson_tridiag(GridDev g, double2 *__restrict__ spec,
                                                         double inv_norm) {
  // consecutive line groups on one XCD (a permutation: block 0 stays 0)
  poisson_body<PQ>(xcd_tile(blockIdx.x, gridDim.x), g, spec, inv_norm);
}

// Grouped: blockIdx.y = instance.
template <int PQ>
__global__ __launch_bounds__(512) void k_poisson_tridiag_g(const StepDesc *__restrict__ tab) {
  const StepDesc &d = tab[blockIdx.y];
  if (inst_stopped(d)) return;
  poisson_body<PQ>(blockIdx.x, glob(d.g), glob(d.spec), d.inv_norm);
}

// Launch shape of k_poisson_tridiag for a line of n rows: PT partitions of at
// most PQ rows (and at least 2), LB lines per block with LB*PT <= 256
// (TC_POISSON_MAXT, at most 512).  Smaller and larger blocks measured slower
// for both row counts: at 512x512x385, 13 rows, 351 us against 452 (128) and
// 594 (512); at 1024x1024x769, 26 rows, 3261 us against 4267 (128) and 3982
// (512).
// Dynamic LDS of k_poisson_tridiag: the padded reduced systems of lb lines
// plus the gauge line (2n), its multipliers (n) and four coefficient tables.
size_t poisson_lds_doubles(int pt, int lb, int n) {
  return (size_t)lb * (2 * (size_t)pt * 4 + 4) + 7 * (size_t)n;
}

void poisson_part_shape(int n, int PQ, int &pt, int &lb) {
  pt = (n + PQ - 1) / PQ;
  if (pt < 1) pt = 1;
  while (pt > 1 && n < 2 * pt) pt--;
  static const int env_maxt = [] {
    const char *e = std::getenv("TC_POISSON_MAXT");
    return e ? std::atoi(e) : 0;
  }();
  const int maxt = env_maxt > 0 ? std::min(env_maxt, 512) : 256;
  lb = 64;
  while (lb > 1 && lb * pt > maxt) lb /= 2;
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

// LDS of one block: pl.lines lines of pl.n points plus the padding slots
static size_t fft_lds_bytes(const FFTPlan &pl) {
  const long long last = (long long)pl.lines * pl.n - 1;
  return (size_t)(last + (last >> pl.padsh) + 1) * sizeof(double2);
}

template <int ME>
static void launch_theta(tc_ctx *c, bool fwd, long long nrows, unsigned blocks, size_t lds,
                         const DivIn *dv) {
  if (fwd && dv && c->divfft_stage)
    hipLaunchKernelGGL((k_div_fft_theta_fwd<ME, true>), dim3(blocks), dim3(256), lds, c->stream,
                       *dv, c->gd, c->spec, c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
  else if (fwd && dv)
    hipLaunchKernelGGL((k_div_fft_theta_fwd<ME, false>), dim3(blocks), dim3(256), lds, c->stream,
                       *dv, c->gd, c->spec, c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
  else if (fwd)
    hipLaunchKernelGGL(k_fft_theta_fwd<ME>, dim3(blocks), dim3(256), lds, c->stream, c->f[TC_DPH],
                       c->spec, c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
  else
    hipLaunchKernelGGL(k_fft_theta_inv<ME>, dim3(blocks), dim3(256), lds, c->stream, c->spec,
                       c->f[TC_DPH], c->plan_t, c->tw_t, nrows, c->nth2, c->sstr);
}

static void theta_pass(tc_ctx *c, bool fwd, const DivIn *dv = nullptr) {
  const long long nrows = (long long)c->nrm * c->nzm;
  const long long npairs = (nrows + 1) / 2;
  const FFTPlan &pt = c->plan_t;
  const size_t lds = fft_lds_bytes(pt);
  const unsigned blocks = (unsigned)((npairs + pt.lines - 1) / pt.lines);
  switch (pt.me) {
    case 4: launch_theta<4>(c, fwd, nrows, blocks, lds, dv); break;
    case 8: launch_theta<8>(c, fwd, nrows, blocks, lds, dv); break;
    case 16: launch_theta<16>(c, fwd, nrows, blocks, lds, dv); break;
    default: launch_theta<32>(c, fwd, nrows, blocks, lds, dv); break;
  }
}

template <int NT>
static void z_pass_nt(tc_ctx *c, double sg) {
  const FFTPlan &pz = c->plan_z;
  const size_t lds = fft_lds_bytes(pz);
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

void launch_poisson(tc_ctx *c, bool with_divergence, double usdt) {
  const long long M = (long long)c->nzm * c->sstr;
  if (with_divergence) {
    const DivIn dv{c->f[TC_QT], c->f[TC_QR], c->f[TC_QZ], usdt};
    timer_begin(c, "div_fft_theta_fwd");
    theta_pass(c, true, &dv);
  } else {
    timer_begin(c, "fft_theta_fwd");
    theta_pass(c, true);
  }
  timer_end(c);
  timer_begin(c, "fft_z_fwd");
  z_pass(c, -1.0);
  timer_end(c);
  timer_begin(c, "poisson_tridiag");
  {
    int pt, lb;
    const int pq = c->poisson_pq;
    poisson_part_shape(c->nrm, pq, pt, lb);
    const size_t lds = sizeof(double) * poisson_lds_doubles(pt, lb, c->nrm);
    const dim3 grid((unsigned)((M + lb - 1) / lb)), block(lb, pt);
    const double inv_norm = 1.0 / ((double)c->nthm * (double)c->nzm);
    if (pq == 13)
      hipLaunchKernelGGL(k_poisson_tridiag<13>, grid, block, lds, c->stream, c->gd, c->spec, inv_norm);
    else
      hipLaunchKernelGGL(k_poisson_tridiag<26>, grid, block, lds, c->stream, c->gd, c->spec, inv_norm);
  }
  timer_end(c);
  timer_begin(c, "fft_z_inv");
  z_pass(c, 1.0);
  timer_end(c);
  timer_begin(c, "fft_theta_inv");
  theta_pass(c, false);
  timer_end(c);
}

// Grouped Poisson solve of n same-shape instances: divergence fused into the
// forward theta FFT, z FFT, radial solve, inverse z and theta FFTs.
void launch_poisson_g(tc_ctx *c, const StepDesc *tab, int n, hipStream_t st) {
  const FFTPlan &pt = c->plan_t, &pz = c->plan_z;
  const long long nrows = (long long)c->nrm * c->nzm;
  const unsigned tb = (unsigned)(((nrows + 1) / 2 + pt.lines - 1) / pt.lines);
  const size_t lt = fft_lds_bytes(pt), lz = fft_lds_bytes(pz);
  const dim3 gt(tb, (unsigned)n);
#define TC_TG(K, ME) hipLaunchKernelGGL(K<ME>, gt, dim3(256), lt, st, tab)
  switch (pt.me) {
    case 4: TC_TG(k_div_fft_theta_fwd_g, 4); break;
    case 8: TC_TG(k_div_fft_theta_fwd_g, 8); break;
    case 16: TC_TG(k_div_fft_theta_fwd_g, 16); break;
    default: TC_TG(k_div_fft_theta_fwd_g, 32); break;
  }
  const int ntile = (c->nth2 + pz.lines - 1) / pz.lines;
  const int ntiles = ntile * c->nrm;
  for (int pass = 0; pass < 3; pass++) {
    if (pass == 1) {
      int ptn, lb;
      poisson_part_shape(c->nrm, c->poisson_pq, ptn, lb);
      const size_t lds = sizeof(double) * poisson_lds_doubles(ptn, lb, c->nrm);
      const long long M = (long long)c->nzm * c->sstr;
      const dim3 grid((unsigned)((M + lb - 1) / lb), (unsigned)n), block(lb, ptn);
      if (c->poisson_pq == 13)
        hipLaunchKernelGGL(k_poisson_tridiag_g<13>, grid, block, lds, st, tab);
      else
        hipLaunchKernelGGL(k_poisson_tridiag_g<26>, grid, block, lds, st, tab);
      continue;
    }
    const double sg = pass == 0 ? -1.0 : 1.0;
    const dim3 gz((unsigned)ntiles, (unsigned)n);
#define TC_ZG(ME, NT) hipLaunchKernelGGL((k_fft_z_g<ME, NT>), gz, dim3(NT), lz, st, tab, ntiles, ntile, sg)
    if (pz.nt == 512) {
      switch (pz.me) {
        case 4: TC_ZG(4, 512); break;
        case 8: TC_ZG(8, 512); break;
        case 16: TC_ZG(16, 512); break;
        default: TC_ZG(32, 512); break;
      }
    } else {
      switch (pz.me) {
        case 4: TC_ZG(4, 256); break;
        case 8: TC_ZG(8, 256); break;
        case 16: TC_ZG(16, 256); break;
        default: TC_ZG(32, 256); break;
      }
    }
#undef TC_ZG
  }
  switch (pt.me) {
    case 4: TC_TG(k_fft_theta_inv_g, 4); break;
    case 8: TC_TG(k_fft_theta_inv_g, 8); break;
    case 16: TC_TG(k_fft_theta_inv_g, 16); break;
    default: TC_TG(k_fft_theta_inv_g, 32); break;
  }
#undef TC_TG
}

}  // namespace tc
