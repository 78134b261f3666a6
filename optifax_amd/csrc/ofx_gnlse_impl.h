// Body of the fibre GNLSE solver: one workgroup integrates one fibre
// instance over z (gnlsefib.m:1-89), the field held in LDS for the FFTs.
//
// The ODE is gnlsefib.m's interaction-picture form, dAW/dz = rhs(z, AW)
// (gnlsefib.m:49-61), with MATLAB's fft / ifft conventions:
//   AT = fft(AW exp(L z)),  IT = |AT|^2,
//   RS = dT fr fft(ifft(IT) RW)                        (Raman, when given)
//   M  = ifft(AT ((1 - fr) IT + RS))   or  ifft(AT IT)  (no Raman)
//   rhs = i gamma W M exp(-L z),  L = i B - alpha / 2
// per evaluation two length-n FFTs and, with Raman, two length-n/2 FFTs
// (the real convolution as half-length transforms of packed data), each a
// radix-4 Stockham transform in LDS (fft_fused) over the workgroup's two LDS
// buffers with the pointwise steps in its first and last passes, twiddles
// from per-pass tables in LDS.
//
// Integration: the Dormand-Prince 5(4) pair with the step control of
// MATLAB's ode45 at NormControl on (Shampine & Reichelt 1997), outputs at
// the requested points through the pair's quartic continuous extension.
// Every thread runs the same scalar control from the same FP64 block sums
// (fixed summation order), so all take the same decisions without a
// broadcast.  Stage vectors f_1..f_7, y, ynew and the stage input live in
// global memory (per instance 11 n complex values, L2-resident for the
// sizes an LDS field allows).
//
// Compiles as HIP device code (ofx_gnlse.hip) and, with OFX_EMU, as plain
// C++ where a workgroup is one CPU thread (tests only).
#pragma once
#include <vector>

#include "ofx_pass_impl.h"

namespace ofx {

struct GnlseArgs {
  int n, lg, nsaves, n_inst;
  const double* zs;      // [nsaves] requested points, zs[0] = start
  const double* B;       // [n_inst][n] Im L = the dispersion B(V), FFT order
  const double* lre;     // [n_inst] Re L = -alpha / 2 (the same at every frequency)
  const void* W;         // [n_inst][n] R: V + w0 or 1, FFT order
  const void* RW;        // [n_inst][n] cpx<R>, or null: no Raman
  const double* gamma;   // [n_inst] (gamma / w0 when self-steepening is on)
  const double* fr;      // [n_inst] Raman fraction
  const double* dT;      // [n_inst] time step
  double rtol, atol;
  long long max_steps;   // accepted + rejected steps before the instance gives up
  void* ws;              // [n_inst][11 n] cpx<R> work vectors
  const void* tw;        // [fft_table_size(lg)] cpx<R>: the FFT passes' twiddles
  const double* y0;      // [n_inst][n] complex: AW at zs[0]
  double* out;           // [n_inst][nsaves][n] complex: AW at each requested point
  long long* stats;      // [n_inst][4]: accepted, rejected, evaluations, status
};

enum { GNLSE_OK = 0, GNLSE_HMIN = 1, GNLSE_CAP = 2, GNLSE_NONFINITE = 3 };

// Dormand-Prince (1980) 5(4) pair; column k of DP_B gives stage k + 2, the
// last column the 5th-order solution; DP_E = b5 - b4; DP_BI the continuous
// extension (Shampine 1986): y(t + s h) = y + h sum_j f_j sum_m DP_BI[j][m] s^(m+1).
// (Constant memory on the device: the control indexes them at run time.)
#ifdef OFX_EMU
#define OFX_CONST_TABLE static const
#else
#define OFX_CONST_TABLE static __constant__ const
#endif
OFX_CONST_TABLE double DP_A[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
OFX_CONST_TABLE double DP_B[7][6] = {
    {1.0 / 5, 3.0 / 40, 44.0 / 45, 19372.0 / 6561, 9017.0 / 3168, 35.0 / 384},
    {0, 9.0 / 40, -56.0 / 15, -25360.0 / 2187, -355.0 / 33, 0},
    {0, 0, 32.0 / 9, 64448.0 / 6561, 46732.0 / 5247, 500.0 / 1113},
    {0, 0, 0, -212.0 / 729, 49.0 / 176, 125.0 / 192},
    {0, 0, 0, 0, -5103.0 / 18656, -2187.0 / 6784},
    {0, 0, 0, 0, 0, 11.0 / 84},
    {0, 0, 0, 0, 0, 0}};
OFX_CONST_TABLE double DP_E[7] = {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920,
                                  -17253.0 / 339200, 22.0 / 525, -1.0 / 40};
OFX_CONST_TABLE double DP_BI[7][4] = {{1, -183.0 / 64, 37.0 / 12, -145.0 / 128},
                                      {0, 0, 0, 0},
                                      {0, 1500.0 / 371, -1000.0 / 159, 1000.0 / 371},
                                      {0, -125.0 / 32, 125.0 / 12, -375.0 / 64},
                                      {0, 9477.0 / 3392, -729.0 / 106, 25515.0 / 6784},
                                      {0, -11.0 / 7, 11.0 / 3, -55.0 / 28},
                                      {0, 3.0 / 2, -4, 5.0 / 2}};

// Sum of one value per thread over the workgroup in FP64, the same result in
// every thread (xor butterfly within a wave, then the wave partials in wave
// order).  red: one double per wave.
OFX_DEV double wg_sum(double v, double* red) {
#ifdef OFX_EMU
  (void)red;
  return v;
#else
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int nw = (int)(blockDim.x >> 6);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; i++) s += red[i];
  return s;
#endif
}

// MATLAB's eps(t): the spacing of doubles at |t|
OFX_DEV double eps_of(double t) {
  t = fabs(t);
  if (t < 2.2250738585072014e-308) return 4.9406564584124654e-324;
  int e;
  (void)frexp(t, &e);
  return ldexp(1.0, e - 53);
}

// Twiddle tables in LDS.  The radix-4 pass with stride Ns > 1 reads Ns
// entries exp(-2 pi i k / (4 Ns)), k < Ns, at offset (Ns - 4) / 3, for Ns =
// 4, 16, ... up to the largest power of 4 not above n / 4 (consecutive
// butterflies read consecutive entries).  The quarter table qt[m] =
// exp(-2 pi i m / n), m < n / 4, serves the radix-2 pass of odd lengths and
// the real-data step of the Raman convolution: for even lg >= 4 it is the
// last radix-4 table, otherwise it follows the radix-4 tables.
#ifdef OFX_EMU
#define OFX_HD inline
#else
#define OFX_HD __host__ __device__ inline
#endif
OFX_HD int fft_r4_top(int lg) { return 1 << (2 * ((lg - 2) >> 1)); }
OFX_HD bool fft_qt_shared(int lg) { return !(lg & 1) && lg >= 4; }
OFX_HD int fft_qt_offset(int lg) {
  const int top = fft_r4_top(lg);
  return fft_qt_shared(lg) ? (top - 4) / 3 : (4 * top - 4) / 3;
}
OFX_HD int fft_table_size(int lg) {
  return (4 * fft_r4_top(lg) - 4) / 3 + (fft_qt_shared(lg) ? 0 : (1 << lg) / 4);
}

template <typename R>
std::vector<cpx<R>> fft_twiddles(int lg) {
  const int n = 1 << lg, top = fft_r4_top(lg);
  std::vector<cpx<R>> t((size_t)fft_table_size(lg));
  auto put = [&](int at, double ph) {
    t[(size_t)at].x = (R)cos(ph);
    t[(size_t)at].y = (R)sin(ph);
  };
  for (int ns = 4; ns <= top; ns <<= 2)
    for (int k = 0; k < ns; k++) put((ns - 4) / 3 + k, -2.0 * M_PI * (double)k / (4.0 * ns));
  if (!fft_qt_shared(lg))
    for (int m = 0; m < n / 4; m++) put(fft_qt_offset(lg) + m, -2.0 * M_PI * (double)m / n);
  return t;
}

// exp(-2 pi i k / L) for k < L / 2 from the quarter table of the length n = L s
template <typename R>
OFX_DEV cpx<R> tw_half(const cpx<R>* qt, int n, int s, int k) {
  const int m = k * s, q = n >> 2;
  if (m < q) return qt[m];
  const cpx<R> v = qt[m - q];
  return mkc<R>(v.y, -v.x);   // -i v
}

// Radix-4 butterfly on (v0, v1, v2, v3) in place, the inputs v1..v3 first
// multiplied by w1, w1^2, w1^3 (conjugated for the inverse) when TW.
template <typename R, int SIGN, bool TW>
OFX_DEV void bfly4(cpx<R>& v0, cpx<R>& v1, cpx<R>& v2, cpx<R>& v3, cpx<R> w1) {
  typedef cpx<R> C;
  if (TW) {
    const C w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    if (SIGN > 0) {
      v1 = cmulc(v1, w1);
      v2 = cmulc(v2, w2);
      v3 = cmulc(v3, w3);
    } else {
      v1 = cmul(v1, w1);
      v2 = cmul(v2, w2);
      v3 = cmul(v3, w3);
    }
  }
  const C s02 = cadd(v0, v2), d02 = csub(v0, v2), s13 = cadd(v1, v3), d13 = csub(v1, v3);
  const C rd = (SIGN < 0) ? mkc<R>(d13.y, -d13.x) : mkc<R>(-d13.y, d13.x);
  v0 = cadd(s02, s13);
  v1 = cadd(d02, rd);
  v2 = csub(s02, s13);
  v3 = csub(d02, rd);
}

// One length-2^lg Stockham FFT in LDS: radix-4 passes (Ns = 1, 4, ...) and,
// for odd lg, a last radix-2 pass (Ns = L / 2, twiddles from the quarter
// table qt of the length L qs).  SIGN = -1 forward, +1 inverse, both
// unnormalised.  The pointwise work on either side is fused into the first
// and last passes: ld(i) gives input element i, st(o, i, v) takes output
// element i, with o the LDS buffer the last pass may write (returned).  Pass
// p writes b[p & 1]; ld may read LDS only in b1.  Every pass ends in a
// barrier.
template <typename R, int SIGN, typename Ld, typename St>
OFX_DEV cpx<R>* fft_fused(int lg, cpx<R>* b0, cpx<R>* b1, const cpx<R>* tw, const cpx<R>* qt,
                          int qs, Ld ld, St st) {
  typedef cpx<R> C;
  const int L = 1 << lg, q4 = L >> 2, h = L >> 1;
  const int tid = OFX_TID, nt = OFX_NT;
  const int n4 = lg >> 1, np = n4 + (lg & 1);
  // (buffers chosen by selects, not by indexing an array of pointers: through
  // an array the compiler loses that they are LDS and emits flat accesses)
  C* const outbuf = ((np - 1) & 1) ? b1 : b0;
  const C one = mkc<R>(R(1), R(0));
  if (np == 1) {  // L = 2 or 4: one pass from ld to st
    if (tid == 0) {
      if (lg == 1) {
        const C v0 = ld(0), v1 = ld(1);
        st(outbuf, 0, cadd(v0, v1));
        st(outbuf, 1, csub(v0, v1));
      } else {
        C v0 = ld(0), v1 = ld(1), v2 = ld(2), v3 = ld(3);
        bfly4<R, SIGN, false>(v0, v1, v2, v3, one);
        st(outbuf, 0, v0);
        st(outbuf, 1, v1);
        st(outbuf, 2, v2);
        st(outbuf, 3, v3);
      }
    }
    OFX_SYNC();
    return outbuf;
  }
  // first pass (radix 4, Ns = 1): from ld into b0
#pragma unroll 1
  for (int j = tid; j < q4; j += nt) {
    C v0 = ld(j), v1 = ld(j + q4), v2 = ld(j + 2 * q4), v3 = ld(j + 3 * q4);
    bfly4<R, SIGN, false>(v0, v1, v2, v3, one);
    b0[4 * j] = v0;
    b0[4 * j + 1] = v1;
    b0[4 * j + 2] = v2;
    b0[4 * j + 3] = v3;
  }
  OFX_SYNC();
  // radix-4 passes LDS to LDS (all the others for odd lg, all but the last for even)
  const int mid_end = (lg & 1) ? n4 : n4 - 1;
  int Ns = 4;
  for (int p = 1; p < mid_end; p++, Ns <<= 2) {
    const C* in = (p & 1) ? b0 : b1;
    C* o = (p & 1) ? b1 : b0;
    const C* tws = tw + (Ns - 4) / 3;
#pragma unroll 2
    for (int j = tid; j < q4; j += nt) {
      const int k = j & (Ns - 1);
      C v0 = in[j], v1 = in[j + q4], v2 = in[j + 2 * q4], v3 = in[j + 3 * q4];
      bfly4<R, SIGN, true>(v0, v1, v2, v3, tws[k]);   // tws[k] = exp(-2 pi i k / (4 Ns))
      const int base = ((j - k) << 2) + k;
      o[base] = v0;
      o[base + Ns] = v1;
      o[base + 2 * Ns] = v2;
      o[base + 3 * Ns] = v3;
    }
    OFX_SYNC();
  }
  // last pass: element i of the result to st
  const C* in = (np & 1) ? b1 : b0;   // written by pass np - 2
  if (lg & 1) {  // radix 2, Ns = L / 2
#pragma unroll 1
    for (int j = tid; j < h; j += nt) {
      const C w = tw_half(qt, L * qs, qs, j);   // exp(-2 pi i j / L)
      const C v0 = in[j], v1 = SIGN > 0 ? cmulc(in[j + h], w) : cmul(in[j + h], w);
      st(outbuf, j, cadd(v0, v1));
      st(outbuf, j + h, csub(v0, v1));
    }
  } else {  // radix 4, Ns = L / 4
    const C* tws = tw + (q4 - 4) / 3;
#pragma unroll 1
    for (int j = tid; j < q4; j += nt) {
      C v0 = in[j], v1 = in[j + q4], v2 = in[j + 2 * q4], v3 = in[j + 3 * q4];
      bfly4<R, SIGN, true>(v0, v1, v2, v3, tws[j]);
      st(outbuf, j, v0);
      st(outbuf, j + q4, v1);
      st(outbuf, j + 2 * q4, v2);
      st(outbuf, j + 3 * q4, v3);
    }
  }
  OFX_SYNC();
  return outbuf;
}

template <typename R>
struct GnlseInst {
  const GnlseArgs* a;
  int n, lg;
  cpx<R>* lds0;
  cpx<R>* lds1;
  double* red;
  const double* B;
  double lre;
  const R* W;
  const cpx<R>* RW;
  const cpx<R>* tw;  // radix-4 twiddle tables (LDS)
  const cpx<R>* qt;  // quarter table exp(-2 pi i m / n), m < n / 4 (LDS)
  double gamma, fr, dT;
  cpx<R>* at;  // AT of the current evaluation
};

// exp(+-i B_i z): the phase formed and reduced in FP64
template <typename R>
OFX_DEV cpx<R> phase_factor(const OFX_GLOBAL double* B, int i, double z) {
  return expi(B[i] * z, R(0));
}

// f = rhs(z, y) (gnlsefib.m:49-61); y, f in global memory.  The pointwise
// steps run inside the first and last passes of the FFTs around them.
//
// The Raman term fft(ifft(IT) RW) is a circular convolution of two real
// sequences (IT = |AT|^2, and RW the spectrum of the real response), so its
// result is real, and both of its transforms run at half length on the
// packed sequence z[m] = IT[2m] + i IT[2m+1]:
//   Zi = sum_m z[m] e^{+2 pi i m k / h}, h = n / 2 (one inverse transform)
//   E = (Zi[k] + conj Zi[h-k]) / 2,  O = (Zi[k] - conj Zi[h-k]) / 2i
//   U[k] = E + e^{2 pi i k / n} O,  U[k+h] = E - e^{2 pi i k / n} O   (= n ifft(IT))
//   X = U RW / n;  P = X[k] + X[k+h],  Q = (X[k] - X[k+h]) e^{-2 pi i k / n}
//   sum_k (P + i Q) e^{-2 pi i m k / h} = RS'[2m] + i RS'[2m+1]  (one forward transform)
// with RS' = fft(X).  The real part is the whole result: the imaginary part
// a full-length complex evaluation carries is rounding.
template <typename R>
OFX_DEV void gnlse_rhs(const GnlseInst<R>& c, double z, const cpx<R>* y, cpx<R>* f) {
  typedef cpx<R> C;
  const int lg = c.lg, n = c.n, h = n >> 1;
  const R inv_n = (R)(1.0 / (double)n);
  // exp(L z) = exp(Re L z) exp(i B z); Re L is one number per fibre
  const R gf = (R)exp(c.lre * z), gb = (R)(exp(-c.lre * z) / (double)n);
  const OFX_GLOBAL double* B = gl(c.B);
  OFX_GLOBAL C* const at = gl(c.at);
  const C* qt = c.qt;
  const OFX_GLOBAL C* yg = gl(y);
  OFX_GLOBAL C* fo = gl(f);
  // AT = fft(AW exp(L z))
  auto ld_y = [=](int i) { return cscale(cmul(ldg(yg, i), phase_factor<R>(B, i, z)), gf); };
  // rhs = i gamma W (M / n) exp(-L z), M = n ifft(.)
  const R gW = (R)c.gamma;
  const OFX_GLOBAL R* W = gl(c.W);
  auto st_f = [=](C*, int i, C v) {
    const C m = cscale(cmulc(v, phase_factor<R>(B, i, z)), gb);
    const R s = gW * W[i];
    stg(fo, i, mkc<R>(-s * m.y, s * m.x));
  };
  C* X;
  if (c.RW) {
    // AT kept for the last product; IT = |AT|^2 as n reals in the low half of X
    X = fft_fused<R, -1>(lg, c.lds0, c.lds1, c.tw, qt, 1, ld_y, [=](C* o, int i, C v) {
      stg(at, i, v);
      reinterpret_cast<R*>(o)[i] = v.x * v.x + v.y * v.y;
    });
    C* const Y = X == c.lds0 ? c.lds1 : c.lds0;
    // Zi: half-length inverse transform of z within X
    C* const Xz = X;
    C* Zi = fft_fused<R, 1>(lg - 1, X + h, X, c.tw, qt, 2, [=](int m) { return Xz[m]; },
                            [](C* o, int m, C v) { o[m] = v; });
    C* const Zp = Zi == X ? X + h : X;
    const OFX_GLOBAL C* RW = gl(c.RW);
    const int tid = OFX_TID, nt = OFX_NT;
    for (int k = tid; k < h; k += nt) {
      const C a = Zi[k], b = Zi[k == 0 ? 0 : h - k];
      const C bc = mkc<R>(b.x, -b.y);
      const C E = cscale(cadd(a, bc), R(0.5));
      const C d = csub(a, bc);
      const C O = mkc<R>(R(0.5) * d.y, R(-0.5) * d.x);   // d / 2i
      const C w = tw_half(qt, n, 1, k);                    // exp(-2 pi i k / n)
      const C wO = cmulc(O, w);
      const C x0 = cscale(cmul(cadd(E, wO), ldg(RW, k)), inv_n);
      const C x1 = cscale(cmul(csub(E, wO), ldg(RW, k + h)), inv_n);
      const C P = cadd(x0, x1), Q = cmul(csub(x0, x1), w);
      Zp[k] = mkc<R>(P.x - Q.y, P.y + Q.x);   // P + i Q
    }
    OFX_SYNC();
    // RS' by the half-length forward transform; into Y: AT ((1 - fr) IT + dT fr RS')
    const R dfr = (R)(c.dT * c.fr), one_fr = (R)(1.0 - c.fr);
    (void)fft_fused<R, -1>(lg - 1, Zi, Zp, c.tw, qt, 2, [=](int m) { return Zp[m]; },
                           [=](C*, int m, C v) {
                             const C a0 = ldg(at, 2 * m), a1 = ldg(at, 2 * m + 1);
                             Y[2 * m] = cscale(a0, one_fr * (a0.x * a0.x + a0.y * a0.y) + dfr * v.x);
                             Y[2 * m + 1] =
                                 cscale(a1, one_fr * (a1.x * a1.x + a1.y * a1.y) + dfr * v.y);
                           });
    X = Y;
  } else {
    // AT |AT|^2
    X = fft_fused<R, -1>(lg, c.lds0, c.lds1, c.tw, qt, 1, ld_y,
                         [=](C* o, int i, C v) { o[i] = cscale(v, v.x * v.x + v.y * v.y); });
  }
  const C* Xm = X;
  (void)fft_fused<R, 1>(lg, X == c.lds0 ? c.lds1 : c.lds0, X, c.tw, qt, 1,
                        [=](int i) { return Xm[i]; }, st_f);
}

// dst = y + sum_j cf[j] f_j over the first nj stages, in stage order.  Every
// stage vector is read (those past nj weigh nothing, by a select: they may
// hold anything) so the loads of one point issue together; a zero
// coefficient adds a zero, which leaves a finite sum unchanged.
template <typename R>
OFX_DEV void stage_comb(int n, const cpx<R>* y, cpx<R>* const* f, const double* cf, int nj,
                        cpx<R>* dst) {
  R c[7];
  const OFX_GLOBAL cpx<R>* fg[7];
#pragma unroll
  for (int j = 0; j < 7; j++) {
    c[j] = j < nj ? (R)cf[j] : R(0);
    fg[j] = gl(f[j]);
  }
  const OFX_GLOBAL cpx<R>* yg = gl(y);
  OFX_GLOBAL cpx<R>* dg = gl(dst);
#pragma unroll 2
  for (int i = OFX_TID; i < n; i += OFX_NT) {
    R ax = R(0), ay = R(0);
#pragma unroll
    for (int j = 0; j < 7; j++) {
      const R vx = fg[j][i].x, vy = fg[j][i].y;
      ax = j < nj ? ax + c[j] * vx : ax;
      ay = j < nj ? ay + c[j] * vy : ay;
    }
    dg[i].x = yg[i].x + ax;
    dg[i].y = yg[i].y + ay;
  }
}

template <typename R>
OFX_DEV double norm2(int n, const cpx<R>* v, double* red) {
  const OFX_GLOBAL cpx<R>* vg = gl(v);
  double s = 0.0;
#pragma unroll 4
  for (int i = OFX_TID; i < n; i += OFX_NT) {
    const double x = (double)vg[i].x, y = (double)vg[i].y;
    s += x * x + y * y;
  }
  return sqrt(wg_sum(s, red));
}

// || sum_j E_j f_j ||, the error weights applied per point in stage order
// (zero weights by a select, as in stage_comb)
template <typename R>
OFX_DEV double err_norm(int n, cpx<R>* const* f, const double* E, double* red) {
  double e[7];
  const OFX_GLOBAL cpx<R>* fg[7];
#pragma unroll
  for (int j = 0; j < 7; j++) {
    e[j] = E[j];
    fg[j] = gl(f[j]);
  }
  double s = 0.0;
#pragma unroll 2
  for (int i = OFX_TID; i < n; i += OFX_NT) {
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int j = 0; j < 7; j++) {
      const double vx = (double)fg[j][i].x, vy = (double)fg[j][i].y;
      ax = e[j] != 0.0 ? ax + e[j] * vx : ax;
      ay = e[j] != 0.0 ? ay + e[j] * vy : ay;
    }
    s += ax * ax + ay * ay;
  }
  return sqrt(wg_sum(s, red));
}

template <typename R>
OFX_DEV void store_out(int n, const cpx<R>* v, double* o) {
  for (int i = OFX_TID; i < n; i += OFX_NT) {
    o[2 * i] = (double)v[i].x;
    o[2 * i + 1] = (double)v[i].y;
  }
}

template <typename R>
OFX_DEV bool all_finite(int n, const cpx<R>* v, double* red) {
  double bad = 0.0;
  for (int i = OFX_TID; i < n; i += OFX_NT)
    if (!(fabs((double)v[i].x) <= 1.7976931348623157e308) ||
        !(fabs((double)v[i].y) <= 1.7976931348623157e308))
      bad = 1.0;
  return wg_sum(bad, red) == 0.0;
}

// Integrate instance `inst` over zs[0] .. zs[nsaves - 1] (ode45, NormControl on).
template <typename R>
OFX_DEV void gnlse_instance(const GnlseArgs& a, int inst, cpx<R>* lds0, cpx<R>* lds1,
                            cpx<R>* ldstw, double* red) {
  typedef cpx<R> C;
  const int n = a.n;
  GnlseInst<R> c;
  c.a = &a;
  c.n = n;
  c.lg = a.lg;
  c.lds0 = lds0;
  c.lds1 = lds1;
  c.red = red;
  c.B = a.B + (size_t)inst * n;
  c.lre = a.lre[inst];
  c.W = (const R*)a.W + (size_t)inst * n;
  c.RW = a.RW ? (const C*)a.RW + (size_t)inst * n : nullptr;
  for (int i = OFX_TID; i < fft_table_size(a.lg); i += OFX_NT) ldstw[i] = ((const C*)a.tw)[i];
  c.tw = ldstw;
  c.qt = ldstw + fft_qt_offset(a.lg);

  c.gamma = a.gamma[inst];
  c.fr = a.fr[inst];
  c.dT = a.dT[inst];
  C* ws = (C*)a.ws + (size_t)inst * 11 * n;
  C* f[7];
  for (int j = 0; j < 7; j++) f[j] = ws + (size_t)j * n;
  C* y = ws + (size_t)7 * n;
  C* ynew = ws + (size_t)8 * n;
  C* ys = ws + (size_t)9 * n;
  c.at = ws + (size_t)10 * n;
  double* out = a.out + (size_t)inst * a.nsaves * 2 * n;
  long long* st = a.stats + 4 * (size_t)inst;
  for (int i = OFX_TID; i < n; i += OFX_NT) {
    const double* p = a.y0 + ((size_t)inst * n + i) * 2;
    y[i] = mkc<R>((R)p[0], (R)p[1]);
  }
  OFX_SYNC();
  store_out(n, y, out);
  const double t0 = a.zs[0], tfinal = a.zs[a.nsaves - 1];
  const double tdir = tfinal >= t0 ? 1.0 : -1.0;
  const double rtol = a.rtol, pw = 1.0 / 5.0;
  const double threshold = a.atol / rtol;
  const double hmax = 0.1 * fabs(tfinal - t0);
  double t = t0;
  gnlse_rhs(c, t, y, f[0]);
  long long nfev = 1, acc = 0, rej = 0;
  int status = GNLSE_OK;
  double normy = norm2(n, y, red);
  double hmin = 16.0 * eps_of(t);
  double absh = fmin(hmax, fabs(a.zs[1] - a.zs[0]));
  const double rh = (norm2(n, f[0], red) / fmax(normy, threshold)) / (0.8 * pow(rtol, pw));
  if (absh * rh > 1.0) absh = 1.0 / rh;
  absh = fmax(absh, hmin);
  int nxt = 1;
  bool done = false;
  while (!done) {
    hmin = 16.0 * eps_of(t);
    absh = fmin(hmax, fmax(hmin, absh));
    double h = tdir * absh;
    if (1.1 * absh >= fabs(tfinal - t)) {
      h = tfinal - t;
      absh = fabs(h);
      done = true;
    }
    bool nofailed = true;
    double err = 0.0, normynew = 0.0, tnew = t;
    for (;;) {
      // stages 2..7: one copy of the right-hand side in the code
#pragma unroll 1
      for (int k = 0; k < 6; k++) {
        const double hk = h;
        double tk;
        C* dst;
        if (k < 5) {
          tk = t + h * DP_A[k];
          dst = ys;
        } else {
          tnew = t + h * DP_A[5];
          if (done) tnew = tfinal;
          h = tnew - t;
          tk = tnew;
          dst = ynew;
        }
        double cf[6];
#pragma unroll
        for (int j = 0; j < 6; j++) cf[j] = j <= k ? hk * DP_B[j][k] : 0.0;
        stage_comb(n, y, f, cf, k + 1, dst);
        OFX_SYNC();
        gnlse_rhs(c, tk, dst, k < 5 ? ws + (size_t)(k + 1) * n : f[6]);
      }
      nfev += 6;
      normynew = norm2(n, ynew, red);
      const double errwt = fmax(fmax(normy, normynew), threshold);
      err = absh * (err_norm(n, f, DP_E, red) / errwt);
      if (!(err == err)) {  // NaN: the field is not finite
        status = GNLSE_NONFINITE;
        break;
      }
      if (acc + rej + 1 > a.max_steps) {
        status = GNLSE_CAP;
        break;
      }
      if (err > rtol) {
        rej++;
        if (absh <= hmin) {
          status = GNLSE_HMIN;
          break;
        }
        if (nofailed) {
          nofailed = false;
          absh = fmax(hmin, absh * fmax(0.1, 0.8 * pow(rtol / err, pw)));
        } else {
          absh = fmax(hmin, 0.5 * absh);
        }
        h = tdir * absh;
        done = false;
      } else {
        break;
      }
    }
    if (status != GNLSE_OK) break;
    acc++;
    // requested points inside (t, tnew]
    while (nxt < a.nsaves && tdir * (tnew - a.zs[nxt]) >= 0.0) {
      double* o = out + (size_t)nxt * 2 * n;
      if (a.zs[nxt] == tnew) {
        store_out(n, ynew, o);
      } else {
        const double s = (a.zs[nxt] - t) / h;
        const double sv[4] = {s, s * s, s * s * s, s * s * s * s};
        double cs[7];
        for (int j = 0; j < 7; j++)
          cs[j] = h * (((DP_BI[j][0] * sv[0] + DP_BI[j][1] * sv[1]) + DP_BI[j][2] * sv[2]) +
                       DP_BI[j][3] * sv[3]);
        stage_comb(n, y, f, cs, 7, ys);
        store_out(n, ys, o);
      }
      nxt++;
    }
    OFX_SYNC();
    if (done) break;
    if (nofailed) {
      const double temp = 1.25 * pow(err / rtol, pw);
      absh = temp > 0.2 ? absh / temp : 5.0 * absh;
    }
    t = tnew;
    C* sw = y;
    y = ynew;
    ynew = sw;
    normy = normynew;
    sw = f[0];
    f[0] = f[6];
    f[6] = sw;
  }
  if (status == GNLSE_OK && !all_finite(n, y, red)) status = GNLSE_NONFINITE;
  if (OFX_TID == 0) {
    st[0] = acc;
    st[1] = rej;
    st[2] = nfev;
    st[3] = status;
  }
}

// Host side: the device tables from the caller's operators.  L must be
// i B - alpha / 2 (gnlsefib.m:31): its real part the same at every
// frequency.  Returns false when it is not.
template <typename R>
bool gnlse_tables(int n_inst, int n, const double* L, const double* W, const double* RW,
                  std::vector<double>& B, std::vector<double>& lre, std::vector<R>& Wr,
                  std::vector<cpx<R>>& RWr) {
  const size_t nn = (size_t)n;
  B.resize((size_t)n_inst * nn);
  lre.resize(n_inst);
  Wr.resize((size_t)n_inst * nn);
  RWr.resize(RW ? (size_t)n_inst * nn : 0);
  for (int k = 0; k < n_inst; k++) {
    const double* l = L + (size_t)k * 2 * nn;
    lre[k] = l[0];
    for (size_t i = 0; i < nn; i++) {
      if (l[2 * i] != lre[k]) return false;
      B[(size_t)k * nn + i] = l[2 * i + 1];
      Wr[(size_t)k * nn + i] = (R)W[(size_t)k * nn + i];
      if (RW) {
        RWr[(size_t)k * nn + i].x = (R)RW[((size_t)k * nn + i) * 2];
        RWr[(size_t)k * nn + i].y = (R)RW[((size_t)k * nn + i) * 2 + 1];
      }
    }
  }
  return true;
}

}  // namespace ofx
