// Accuracy of the kernels' exp(i x) from an FP64 argument (ofx_pass_impl.h,
// compiled as plain C++ with OFX_EMU): prints the largest |expi(x) - e^{ix}|
// over a sample of x (up to 1e10) for the FP32 and the FP64 forms, e^{ix}
// from long double cosl / sinl.  Test infrastructure; never part of the product.
#include <math.h>
#include <stdio.h>

#include "../../optifax_amd/csrc/ofx_pass_impl.h"

int main() {
  long double e32 = 0.0L, e64 = 0.0L;
  unsigned long long s = 88172645463325252ULL;   // xorshift64
  for (long i = 0; i < 4000000; i++) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const double u = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    // small, large and (one in 16) very large arguments, beyond an int's quadrant count
    const double x = (i % 16 == 1) ? u * 2.0e10 : ((i % 4 == 0) ? u * 8.0 : u * 2.0e4);
    const long double c = cosl((long double)x), sn = sinl((long double)x);
    const ofx::cpx<float> f = ofx::expi(x, 0.0f);
    const ofx::cpx<double> d = ofx::expi(x, 0.0);
    const long double ef = hypotl((long double)f.x - c, (long double)f.y - sn);
    const long double ed = hypotl((long double)d.x - c, (long double)d.y - sn);
    if (ef > e32) e32 = ef;
    if (ed > e64) e64 = ed;
  }
  printf("%.6Le %.6Le\n", e32, e64);
  return 0;
}
