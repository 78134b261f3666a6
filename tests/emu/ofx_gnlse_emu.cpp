// CPU emulation of the fibre GNLSE kernel (test infrastructure; never loaded
// by the product): ofx_gnlse_run with the same arguments as the HIP library,
// each instance's workgroup run by one CPU thread (OFX_EMU: one thread per
// workgroup, LDS as host arrays).
#include <math.h>

#include <thread>
#include <vector>

#include "../../include/optifax_hip.h"
#include "../../optifax_amd/csrc/ofx_gnlse_impl.h"

namespace {

template <typename R>
int run(int n_inst, int n, int lg, int nsaves, const double* zs, const double* y0,
        const double* L, const double* W, const double* RW, const double* gamma, const double* fr,
        const double* dT, double rtol, double atol, long long max_steps, double* out,
        long long* stats) {
  typedef ofx::cpx<R> C;
  const std::vector<C> tw = ofx::fft_twiddles<R>(lg);
  std::vector<double> hB, hlre;
  std::vector<R> hW;
  std::vector<C> hRW;
  if (!ofx::gnlse_tables<R>(n_inst, n, L, W, RW, hB, hlre, hW, hRW)) return OFX_E_ARG;
  std::vector<C> ws((size_t)n_inst * 11 * n);
  ofx::GnlseArgs a;
  a.n = n;
  a.lg = lg;
  a.nsaves = nsaves;
  a.n_inst = n_inst;
  a.zs = zs;
  a.B = hB.data();
  a.lre = hlre.data();
  a.W = hW.data();
  a.RW = RW ? (const void*)hRW.data() : nullptr;
  a.gamma = gamma;
  a.fr = fr;
  a.dT = dT;
  a.rtol = rtol;
  a.atol = atol;
  a.max_steps = max_steps;
  a.ws = ws.data();
  a.tw = tw.data();
  a.y0 = y0;
  a.out = out;
  a.stats = stats;
  std::vector<std::thread> th;
  for (int i = 0; i < n_inst; i++)
    th.emplace_back([&a, &tw, i, n]() {
      std::vector<C> l0(n), l1(n), tq(tw.size() + 1);
      double red[16];
      ofx::gnlse_instance<R>(a, i, l0.data(), l1.data(), tq.data(), red);
    });
  for (auto& t : th) t.join();
  return OFX_OK;
}

}  // namespace

extern "C" int ofx_gnlse_run(int device, int precision, int n_inst, int n, int nsaves,
                             const double* zs, const double* y0, const double* L,
                             const double* W, const double* RW, const double* gamma,
                             const double* fr, const double* dT, double rtol, double atol,
                             long long max_steps, double* out, long long* stats, double* ms) {
  (void)device;
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
    if (!((zs[i] - zs[i - 1]) * dir > 0.0)) return OFX_E_ARG;
  if (ms) *ms = 0.0;
  if (precision == OFX_F64)
    return run<double>(n_inst, n, lg, nsaves, zs, y0, L, W, RW, gamma, fr, dT, rtol, atol,
                       max_steps, out, stats);
  return run<float>(n_inst, n, lg, nsaves, zs, y0, L, W, RW, gamma, fr, dT, rtol, atol, max_steps,
                    out, stats);
}
