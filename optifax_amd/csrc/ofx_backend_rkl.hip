// The pass kernel's variant for launches whose instances all use RK23: the
// RK23 pass reads the cluster's descriptor from a copy in LDS (OFX_RKL in
// ofx_pass_impl.h).  It is a translation unit of its own so that its code
// does not share a code object with the default kernel, whose register
// allocation the variant moved when compiled beside it.
#define OFX_RKL 1
#include <hip/hip_runtime.h>

#include "../../include/optifax_hip.h"
#include "ofx_backend.h"
#include "ofx_pass_impl.h"

namespace ofx {

template <typename R, int XM>
__global__ void __launch_bounds__(512) pass_kernel_rkl(LaunchArgs a) {
  extern __shared__ __align__(16) unsigned char ofx_lds_raw[];
  __shared__ double red[64];
  __shared__ int sflag[4];
  __shared__ ClusterDesc cdl;
  cluster_main<R, XM>(a, (int)blockIdx.x, reinterpret_cast<cpx<R>*>(ofx_lds_raw), red, sflag, &cdl);
}

const void* rkl_kernel(int precision, int xmode) {
  const bool bar = (xmode & 1) == XM_BARRIER;
  if (precision == OFX_F64)
    return bar ? (const void*)pass_kernel_rkl<double, XM_BARRIER>
               : (const void*)pass_kernel_rkl<double, XM_TAGS>;
  return bar ? (const void*)pass_kernel_rkl<float, XM_BARRIER>
             : (const void*)pass_kernel_rkl<float, XM_TAGS>;
}

}  // namespace ofx
