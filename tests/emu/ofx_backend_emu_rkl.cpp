// CPU emulation of the pass kernel's RK23 variant (TEST INFRASTRUCTURE):
// ofx_pass_impl.h compiled with OFX_EMU and OFX_RKL, the RK23 pass reading
// the cluster's descriptor from a per-workgroup copy, in a translation unit
// of its own as on the device (optifax_amd/csrc/ofx_backend_rkl.hip).
#define OFX_RKL 1
#include <stddef.h>

#include <vector>

#include "../../include/optifax_hip.h"
#include "../../optifax_amd/csrc/ofx_backend.h"
#include "../../optifax_amd/csrc/ofx_pass_impl.h"

namespace ofx {

template <typename R, int XM>
static void run_wg_rkl(const LaunchArgs* a, int wg, size_t lds) {
  std::vector<unsigned char> mem(lds + 64);
  double red[64];
  int sflag[4];
  ClusterDesc cdl;
  cluster_main<R, XM>(*a, wg, reinterpret_cast<cpx<R>*>(mem.data()), red, sflag, &cdl);
}

// the workgroup body of the variant for a precision and exchange mode
void (*emu_rkl_fn(int precision, int xmode))(const LaunchArgs*, int, size_t) {
  const bool bar = (xmode & 1) == XM_BARRIER;
  if (precision == OFX_F64) return bar ? run_wg_rkl<double, XM_BARRIER> : run_wg_rkl<double, XM_TAGS>;
  return bar ? run_wg_rkl<float, XM_BARRIER> : run_wg_rkl<float, XM_TAGS>;
}

}  // namespace ofx
