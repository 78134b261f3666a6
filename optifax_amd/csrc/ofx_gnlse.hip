// Fibre GNLSE solver on gfx950 (gnlsefib.m): one workgroup per fibre
// instance, the whole field in LDS (two buffers of n complex samples for the
// Stockham FFTs and up to 0.42 n twiddles: n <= 2^13 in FP32, 2^12 in FP64
// within the 160 KB of LDS),
// Dormand-Prince 5(4) with ode45's step control, outputs at the requested
// points.  The body is ofx_gnlse_impl.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/optifax_hip.h"
#include "ofx_gnlse_impl.h"

namespace ofx {

template <typename R>
__global__ void __launch_bounds__(512) gnlse_kernel(GnlseArgs a) {
  extern __shared__ __align__(16) unsigned char gnlse_smem[];
  cpx<R>* l0 = reinterpret_cast<cpx<R>*>(gnlse_smem);
  cpx<R>* l1 = l0 + a.n;
  cpx<R>* tw = l1 + a.n;
  double* red = reinterpret_cast<double*>(tw + fft_table_size(a.lg));
  gnlse_instance<R>(a, blockIdx.x, l0, l1, tw, red);
}

}  // namespace ofx

namespace {

struct DBuf {
  void* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

int hip_ok(hipError_t e) { return e == hipSuccess ? OFX_OK : OFX_E_HIP; }

template <typename T>
int upload(DBuf& b, const T* src, size_t n) {
  if (hipMalloc(&b.p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return OFX_E_NOMEM;
  if (n && src) return hip_ok(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return OFX_OK;
}

template <typename R>
int gnlse_run(int n_inst, int n, int lg, int nsaves, const double* zs, const double* y0,
              const double* L, const double* W, const double* RW, const double* gamma,
              const double* fr, const double* dT, double rtol, double atol, long long max_steps,
              double* out, long long* stats, double* ms) {
  typedef ofx::cpx<R> C;
  // the FFT passes' twiddles, built in FP64 and rounded once
  const std::vector<C> tw = ofx::fft_twiddles<R>(lg);
  const size_t ni = (size_t)n_inst, nn = (size_t)n;
  std::vector<double> hB, hlre;
  std::vector<R> hW;
  std::vector<C> hRW;
  if (!ofx::gnlse_tables<R>(n_inst, n, L, W, RW, hB, hlre, hW, hRW)) return OFX_E_ARG;
  DBuf b_zs, b_y0, b_B, b_lre, b_W, b_RW, b_g, b_fr, b_dT, b_tw, b_ws, b_out, b_st;
  int st = upload(b_zs, zs, (size_t)nsaves);
  if (!st) st = upload(b_y0, y0, ni * 2 * nn);
  if (!st) st = upload(b_B, hB.data(), ni * nn);
  if (!st) st = upload(b_lre, hlre.data(), ni);
  if (!st) st = upload(b_W, hW.data(), ni * nn);
  if (!st && RW) st = upload(b_RW, hRW.data(), ni * nn);
  if (!st) st = upload(b_g, gamma, ni);
  if (!st) st = upload(b_fr, fr, ni);
  if (!st) st = upload(b_dT, dT, ni);
  if (!st) st = upload(b_tw, tw.data(), tw.size());
  if (!st) st = upload(b_ws, (const C*)nullptr, ni * 11 * nn);
  if (!st) st = upload(b_out, (const double*)nullptr, ni * (size_t)nsaves * 2 * nn);
  if (!st) st = upload(b_st, (const long long*)nullptr, ni * 4);
  if (st) return st;
  ofx::GnlseArgs a;
  a.n = n;
  a.lg = lg;
  a.nsaves = nsaves;
  a.n_inst = n_inst;
  a.zs = (const double*)b_zs.p;
  a.B = (const double*)b_B.p;
  a.lre = (const double*)b_lre.p;
  a.W = b_W.p;
  a.RW = RW ? b_RW.p : nullptr;
  a.gamma = (const double*)b_g.p;
  a.fr = (const double*)b_fr.p;
  a.dT = (const double*)b_dT.p;
  a.rtol = rtol;
  a.atol = atol;
  a.max_steps = max_steps;
  a.ws = b_ws.p;
  a.tw = b_tw.p;
  a.y0 = (const double*)b_y0.p;
  a.out = (double*)b_out.p;
  a.stats = (long long*)b_st.p;
  // threads: a quarter of the points (one radix-4 butterfly each), 64..512
  const int nt = std::min(512, std::max(64, n / 4));
  const size_t lds = (2 * nn + tw.size()) * sizeof(C) + 8 * (size_t)(nt / 64);
  st = hip_ok(hipFuncSetAttribute((const void*)ofx::gnlse_kernel<R>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (st) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st) st = hip_ok(hipEventCreate(&e0));
  if (!st) st = hip_ok(hipEventCreate(&e1));
  if (!st) st = hip_ok(hipEventRecord(e0, 0));
  if (!st) {
    hipLaunchKernelGGL(ofx::gnlse_kernel<R>, dim3(n_inst), dim3(nt), lds, 0, a);
    st = hip_ok(hipGetLastError());
  }
  if (!st) st = hip_ok(hipEventRecord(e1, 0));
  if (!st) st = hip_ok(hipEventSynchronize(e1));
  if (!st && ms) {
    float f = 0.f;
    st = hip_ok(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
  }
  if (!st) st = hip_ok(hipMemcpy(out, b_out.p, ni * nsaves * 2 * nn * 8, hipMemcpyDeviceToHost));
  if (!st) st = hip_ok(hipMemcpy(stats, b_st.p, ni * 4 * 8, hipMemcpyDeviceToHost));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return st;
}

}  // namespace

extern "C" int ofx_gnlse_run(int device, int precision, int n_inst, int n, int nsaves,
                             const double* zs, const double* y0, const double* L,
                             const double* W, const double* RW, const double* gamma,
                             const double* fr, const double* dT, double rtol, double atol,
                             long long max_steps, double* out, long long* stats, double* ms) {
  int lg = 0;
  while ((1 << lg) < n && lg < 30) lg++;
  const int max_lg = precision == OFX_F64 ? 12 : 13;
  if (n_inst < 1 || n < 4 || (1 << lg) != n || lg > max_lg || nsaves < 3 || !zs || !y0 || !L ||
      !W || !gamma || !fr || !dT || !out || !stats || !(rtol > 0.0) || !(atol > 0.0) ||
      max_steps < 1 || (precision != OFX_F32 && precision != OFX_F64))
    return OFX_E_ARG;
  const double dir = zs[nsaves - 1] - zs[0];
  if (!(dir != 0.0)) return OFX_E_ARG;
  for (int i = 1; i < nsaves; i++)
    if (!((zs[i] - zs[i - 1]) * dir > 0.0)) return OFX_E_ARG;  // strictly monotone
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OFX_E_NODEVICE;
  if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return OFX_E_ARG;
  if (precision == OFX_F64)
    return gnlse_run<double>(n_inst, n, lg, nsaves, zs, y0, L, W, RW, gamma, fr, dT, rtol, atol,
                             max_steps, out, stats, ms);
  return gnlse_run<float>(n_inst, n, lg, nsaves, zs, y0, L, W, RW, gamma, fr, dT, rtol, atol,
                          max_steps, out, stats, ms);
}
