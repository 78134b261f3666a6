// QPM gain maps on gfx950: the two sums behind the reference's GPU phase-
// matching analysis.
//
// 1. qpmgain.m:140-149 evaluates, for every (signal, pump) pair o of the map,
//        As[o] = g[o] * sum_k P(z_k) dz_k exp(i dk[o] z_k)
//    over the crystal's sampled poling P (about 10^6 samples per metre) as a
//    3-D gpuArray [signals x pumps x z] summed over z.  Here one thread owns
//    one output and the z axis arrives as runs: r = (z0, len, weight) stands
//    for `len` samples z0, z0 + dz, ..., with one weight (the poling sign
//    times the sample spacing).  A run's sum is a geometric series,
//        sum_q exp(i dk (z0 + q dz)) = exp(i (dk z0 + th (len-1)/2))
//                                      sin(len th / 2) / sin(th / 2),
//    th = dk dz reduced to (-pi, pi], so a grating of constant domains costs
//    two sines per domain instead of one complex exponential per sample.
//    Runs of one sample give the literal sum for any z.  Phases are formed
//    and reduced mod 2 pi in FP64; everything is FP64.
//
// 2. sincgain.m / sincgain_sparse.m:55-70 evaluate the narrowband mismatch
//    factor gain[g][o] = sum_j (sinc((dk0[o] - m_j kc[g]) L / 2) / m_j)^2 over
//    grating periods g and QPM orders m_j, with MATLAB's normalized sinc
//    sin(pi x)/(pi x).  One thread per (o, g), the argument reduced in FP64.
//
// Both maps are independent per output (no reductions across threads); the
// z runs are staged through LDS in chunks shared by the workgroup.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "../../include/optifax_hip.h"

namespace ofxq {

constexpr int ZS_THREADS = 256;
constexpr int ZS_CHUNK = 512;  // runs staged in LDS at a time (24 KB)

__device__ __forceinline__ double reduce_2pi(double x) {
  // x - 2 pi round(x / 2 pi), two-term (Cody-Waite) split of 2 pi; the fma
  // makes q * TWO_PI_HI exact for |q| < 2^26
  const double TWO_PI_HI = 6.283185307179586232;
  const double TWO_PI_LO = 2.4492935982947064e-16;
  const double q = rint(x * 0.15915494309189533577);
  return fma(-q, TWO_PI_LO, fma(-q, TWO_PI_HI, x));
}

__global__ void __launch_bounds__(ZS_THREADS)
zsum_kernel(int n_out, const double* __restrict__ dk, int n_run, const double* __restrict__ z0,
            const int* __restrict__ len, const double* __restrict__ wgt, double dz,
            double* __restrict__ out) {
  __shared__ double s_z0[ZS_CHUNK];
  __shared__ double s_w[ZS_CHUNK];
  __shared__ double s_m[ZS_CHUNK];
  const int o = blockIdx.x * ZS_THREADS + threadIdx.x;
  const double k = o < n_out ? dk[o] : 0.0;
  // per output: the reduced step phase and the series' denominator
  const double th = reduce_2pi(k * dz);
  const double s_half = sin(0.5 * th);
  double ar = 0.0, ai = 0.0;
  // the series factor depends on the run length only: kept while consecutive
  // runs share it (a uniform grating has one or two lengths)
  double m_prev = -1.0, d_prev = 0.0;
  for (int base = 0; base < n_run; base += ZS_CHUNK) {
    const int cnt = min(ZS_CHUNK, n_run - base);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += ZS_THREADS) {
      s_z0[i] = z0[base + i];
      s_w[i] = wgt[base + i];
      s_m[i] = (double)len[base + i];
    }
    __syncthreads();
    for (int i = 0; i < cnt; i++) {
      const double m = s_m[i];   // the same for every thread
      double d;
      if (m == 1.0) {
        d = 1.0;
      } else if (s_half == 0.0) {
        d = m;  // every sample of the run has the same phase
      } else if (m == m_prev) {
        d = d_prev;
      } else {
        d = sin(reduce_2pi(0.5 * m * th)) / s_half;
        m_prev = m;
        d_prev = d;
      }
      const double ph = reduce_2pi(fma(k, s_z0[i], 0.5 * (m - 1.0) * th));
      double s, c;
      sincos(ph, &s, &c);
      const double a = s_w[i] * d;
      ar = fma(a, c, ar);
      ai = fma(a, s, ai);
    }
  }
  if (o < n_out) {
    out[2 * (size_t)o] = ar;
    out[2 * (size_t)o + 1] = ai;
  }
}

__global__ void sinc_gain_kernel(int n_out, const double* __restrict__ dk0, int n_g,
                                 const double* __restrict__ kc, int n_m,
                                 const double* __restrict__ m, double half_len,
                                 double* __restrict__ gain) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n_out * (size_t)n_g) return;
  const int o = (int)(idx % (size_t)n_out);
  const int g = (int)(idx / (size_t)n_out);
  const double PI = 3.141592653589793116;
  double acc = 0.0;
  for (int j = 0; j < n_m; j++) {
    const double x = (dk0[o] - m[j] * kc[g]) * half_len;  // sinc argument
    double v;
    if (x == 0.0) {
      v = 1.0;
    } else {
      // sin(pi x) with pi x reduced mod 2 pi in FP64
      v = sin(reduce_2pi(PI * x)) / (PI * x);
    }
    v /= m[j];
    acc = fma(v, v, acc);
  }
  gain[idx] = acc;
}

struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) hipFree(p);
  }
};

static int hip_ok(hipError_t e) { return e == hipSuccess ? OFX_OK : OFX_E_HIP; }

template <typename T>
static int up(Buf& b, const T* src, size_t n) {
  if (hipMalloc(&b.p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return OFX_E_NOMEM;
  if (n) return hip_ok(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return OFX_OK;
}

}  // namespace ofxq

using namespace ofxq;

extern "C" int ofx_qpm_zsum(int device, int n_out, const double* dk, int n_run, const double* z0,
                            const int* len, const double* wgt, double dz, double* out,
                            double* ms) {
  if (n_out < 0 || n_run < 0 || !(dz > 0.0) || (n_out && (!dk || !out)) ||
      (n_run && (!z0 || !len || !wgt)))
    return OFX_E_ARG;
  for (int r = 0; r < n_run; r++)
    if (len[r] < 1) return OFX_E_ARG;
  if (n_out == 0) return OFX_OK;
  if (hipSetDevice(device) != hipSuccess) return OFX_E_NODEVICE;
  Buf b_dk, b_z, b_l, b_w, b_o;
  int st = up(b_dk, dk, (size_t)n_out);
  if (!st) st = up(b_z, z0, (size_t)n_run);
  if (!st) st = up(b_l, len, (size_t)n_run);
  if (!st) st = up(b_w, wgt, (size_t)n_run);
  if (!st && hipMalloc(&b_o.p, 16 * (size_t)n_out) != hipSuccess) st = OFX_E_NOMEM;
  if (st) return st;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, 0);
  const int blocks = (n_out + ZS_THREADS - 1) / ZS_THREADS;
  hipLaunchKernelGGL(zsum_kernel, dim3(blocks), dim3(ZS_THREADS), 0, 0, n_out,
                     (const double*)b_dk.p, n_run, (const double*)b_z.p, (const int*)b_l.p,
                     (const double*)b_w.p, dz, (double*)b_o.p);
  st = hip_ok(hipGetLastError());
  hipEventRecord(e1, 0);
  if (!st) st = hip_ok(hipMemcpy(out, b_o.p, 16 * (size_t)n_out, hipMemcpyDeviceToHost));
  if (!st && ms) {
    float f = 0.f;
    hipEventElapsedTime(&f, e0, e1);
    *ms = f;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return st;
}

extern "C" int ofx_sinc_gain(int device, int n_out, const double* dk0, int n_g, const double* kc,
                             int n_m, const double* m, double length, double* gain, double* ms) {
  if (n_out < 0 || n_g < 0 || n_m < 1 || !m || (n_out && (!dk0 || !gain)) || (n_g && !kc))
    return OFX_E_ARG;
  for (int j = 0; j < n_m; j++)
    if (!(m[j] != 0.0)) return OFX_E_ARG;
  const size_t tot = (size_t)n_out * (size_t)n_g;
  if (tot == 0) return OFX_OK;
  if (hipSetDevice(device) != hipSuccess) return OFX_E_NODEVICE;
  Buf b_dk, b_kc, b_m, b_g;
  int st = up(b_dk, dk0, (size_t)n_out);
  if (!st) st = up(b_kc, kc, (size_t)n_g);
  if (!st) st = up(b_m, m, (size_t)n_m);
  if (!st && hipMalloc(&b_g.p, 8 * tot) != hipSuccess) st = OFX_E_NOMEM;
  if (st) return st;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, 0);
  const size_t blocks = (tot + 255) / 256;
  hipLaunchKernelGGL(sinc_gain_kernel, dim3((unsigned)blocks), dim3(256), 0, 0, n_out,
                     (const double*)b_dk.p, n_g, (const double*)b_kc.p, n_m, (const double*)b_m.p,
                     0.5 * length, (double*)b_g.p);
  st = hip_ok(hipGetLastError());
  hipEventRecord(e1, 0);
  if (!st) st = hip_ok(hipMemcpy(gain, b_g.p, 8 * tot, hipMemcpyDeviceToHost));
  if (!st && ms) {
    float f = 0.f;
    hipEventElapsedTime(&f, e0, e1);
    *ms = f;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return st;
}
