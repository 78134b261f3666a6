// OPOmexCavity MEX gateway (no reference counterpart): whole synchronously
// pumped OPO round trips on the device, replacing the MATLAB loop of
// OpticalSim.run (toolbox/lib/classes/OpticalSim.m:230-288) and its per-trip
// OPOmexBatch call.
//
//   [Et, ocSpec, tripLog, status, esdOut, esdX] = OPOmexCavity(Et, alpha_w, G, w0,
//       bdiffw0, h, NSteps, dt, beta_op, maxErr, minErr, sel, pumpSpec, cavMask,
//       ocMask, delay, trips, tol)
//
//   Et       N x B  crystal input of the first trip (fftshifted time); returns
//                   the crystal input of the next trip
//   pumpSpec N x 1  pump spectrum in the crystal frame (FFT order)
//   cavMask  N x 1  crystal output -> next crystal input transfer (FFT order)
//   ocMask   N x 1  crystal output -> output-coupled spectrum (FFT order)
//   delay    1 x B or scalar: cavity detuning per trip [s] (applyGD)
//   trips, tol      round-trip budget; stop an instance when the relative L2
//                   change of its output power spectrum falls below tol
//   ocSpec   N x B  last output-coupled spectrum;  tripLog 4 x trips x B
//                   (conv metric, output energy, attempts, accepted steps)
//   esdOut, esdX    (optional) N x B means over the trips of the output-coupled
//                   and the crystal-output |E(w)|^2, FFT order
#ifdef OFX_MX_SHIM
#include "../../include/optifax_mex.h"
#else
#include "mex.h"
#endif
#include <string.h>

#include <vector>

#include "../../include/optifax_hip.h"

namespace {
ofx_context* g_ctx = nullptr;
void release_context() {
  if (g_ctx) ofx_destroy(g_ctx);
  g_ctx = nullptr;
}

std::vector<double> to_double(const mxArray* a, bool want_complex, const char* name) {
  const size_t n = mxGetNumberOfElements(a);
  std::vector<double> v(want_complex ? 2 * n : n, 0.0);
  const bool cx = mxIsComplex(a);
  if (!want_complex && cx) mexErrMsgIdAndTxt("optifax:OPOmexCavity:type", "%s must be real", name);
  const size_t stride = want_complex ? 2 : 1;
  if (mxGetClassID(a) == mxSINGLE_CLASS) {
    if (cx) {
      const mxComplexSingle* c = mxGetComplexSingles(a);
      for (size_t i = 0; i < n; i++) { v[2 * i] = c[i].real; v[2 * i + 1] = c[i].imag; }
    } else {
      const float* f = mxGetSingles(a);
      for (size_t i = 0; i < n; i++) v[stride * i] = f[i];
    }
  } else if (mxGetClassID(a) == mxDOUBLE_CLASS) {
    if (cx) {
      const mxComplexDouble* c = mxGetComplexDoubles(a);
      for (size_t i = 0; i < n; i++) { v[2 * i] = c[i].real; v[2 * i + 1] = c[i].imag; }
    } else {
      const double* d = mxGetDoubles(a);
      for (size_t i = 0; i < n; i++) v[stride * i] = d[i];
    }
  } else {
    mexErrMsgIdAndTxt("optifax:OPOmexCavity:type", "%s must be single or double", name);
  }
  return v;
}
}  // namespace

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 18) mexErrMsgIdAndTxt("optifax:OPOmexCavity:nrhs", "OPOmexCavity takes 18 inputs");
  const size_t N = mxGetNumberOfElements(prhs[8]);
  const size_t nE = mxGetNumberOfElements(prhs[0]);
  if (N == 0 || nE % N) mexErrMsgIdAndTxt("optifax:OPOmexCavity:size", "numel(Et) must be N*B");
  const int B = (int)(nE / N);
  if (B == 0) mexErrMsgIdAndTxt("optifax:OPOmexCavity:size", "Et holds no instance");
  std::vector<double> E = to_double(prhs[0], true, "Et");
  std::vector<double> alpha = to_double(prhs[1], false, "alpha_w");
  std::vector<double> G = to_double(prhs[2], false, "G");
  std::vector<double> beta = to_double(prhs[8], false, "beta_op");
  std::vector<double> pump = to_double(prhs[12], true, "pumpSpec");
  std::vector<double> cav = to_double(prhs[13], true, "cavMask");
  std::vector<double> ocm = to_double(prhs[14], true, "ocMask");
  std::vector<double> delay = to_double(prhs[15], false, "delay");
  const int trips = (int)mxGetScalar(prhs[16]);
  const double tol = mxGetScalar(prhs[17]);
  if (alpha.size() != N || pump.size() != 2 * N || cav.size() != 2 * N || ocm.size() != 2 * N)
    mexErrMsgIdAndTxt("optifax:OPOmexCavity:size", "tables must have N elements");
  if (delay.size() != 1 && delay.size() != (size_t)B)
    mexErrMsgIdAndTxt("optifax:OPOmexCavity:size", "delay must be scalar or 1 x B");
  if (trips < 1) mexErrMsgIdAndTxt("optifax:OPOmexCavity:trips", "trips must be >= 1");
  ofx_pass_params p;
  memset(&p, 0, sizeof p);
  p.n = (int)N;
  p.w0 = mxGetScalar(prhs[3]);
  p.bdiffw0 = mxGetScalar(prhs[4]);
  p.h = mxGetScalar(prhs[5]);
  p.nsteps = (int)mxGetScalar(prhs[6]);
  p.dt = mxGetScalar(prhs[7]);
  p.max_err = mxGetScalar(prhs[9]);
  p.min_err = mxGetScalar(prhs[10]);
  p.sel = mxGetScalar(prhs[11]);
  p.method = OFX_METHOD_HEUN;
  const bool g_per = G.size() == (size_t)p.nsteps * (size_t)B && B > 1;
  if (!g_per && G.size() != (size_t)p.nsteps)
    mexErrMsgIdAndTxt("optifax:OPOmexCavity:size", "numel(G) must be NSteps or NSteps*B");
  std::vector<ofx_opo_spec> specs((size_t)B);
  for (int b = 0; b < B; b++) {
    ofx_opo_spec& s = specs[b];
    memset(&s, 0, sizeof s);
    s.pass = p;
    s.delay = delay.size() == 1 ? delay[0] : delay[b];
    s.beta = beta.data();
    s.alpha = alpha.data();
    s.G = G.data() + (g_per ? (size_t)b * p.nsteps : 0);
    s.field = E.data() + 2 * N * b;
    s.pump = pump.data();
    s.cav = cav.data();
    s.ocm = ocm.data();
  }
  if (!g_ctx) {
    int st = ofx_create(0, &g_ctx);
    if (st) mexErrMsgIdAndTxt("optifax:OPOmexCavity:device", "no usable HIP device (code %d)", st);
    mexAtExit(release_context);
  }
  ofx_opo* run = nullptr;
  int st = ofx_opo_create(g_ctx, OFX_F32, B, specs.data(), trips, nullptr, &run);
  if (!st) st = ofx_opo_launch(run, trips, tol);
  if (!st) st = ofx_opo_sync(run);
  if (st && st != OFX_E_NONFINITE) {
    if (run) ofx_opo_destroy(run);
    mexErrMsgIdAndTxt("optifax:OPOmexCavity:solver", "%s", ofx_last_error(g_ctx));
  }
  mwSize dE[2] = {(mwSize)N, (mwSize)B};
  plhs[0] = mxCreateNumericArray(2, dE, mxSINGLE_CLASS, mxCOMPLEX);
  plhs[1] = mxCreateNumericArray(2, dE, mxSINGLE_CLASS, mxCOMPLEX);
  mwSize dL[3] = {4, (mwSize)trips, (mwSize)B};
  plhs[2] = mxCreateNumericArray(3, dL, mxDOUBLE_CLASS, mxREAL);
  mwSize dS[2] = {1, (mwSize)B};
  plhs[3] = mxCreateNumericArray(2, dS, mxDOUBLE_CLASS, mxREAL);
  // optional 5th / 6th outputs: the run-averaged output-coupled and
  // crystal-output |E(w)|^2 (N x B, FFT order), ESDOutAverage and the sum
  // behind ESDPumpDepAverage / ESDIdlerICAverage (OpticalSim.m:302-313)
  mwSize dA[2] = {(mwSize)N, (mwSize)B};
  if (nlhs > 4) plhs[4] = mxCreateNumericArray(2, dA, mxDOUBLE_CLASS, mxREAL);
  if (nlhs > 5) plhs[5] = mxCreateNumericArray(2, dA, mxDOUBLE_CLASS, mxREAL);
  std::vector<double> f(2 * N), oc(2 * N);
  for (int b = 0; b < B; b++) {
    int s_b = 0;
    ofx_opo_read(run, b, f.data(), oc.data(), mxGetDoubles(plhs[2]) + (size_t)4 * trips * b, &s_b,
                 nullptr, nullptr, nullptr, nullptr);
    float* pe = (float*)mxGetComplexSingles(plhs[0]) + 2 * N * b;
    float* po = (float*)mxGetComplexSingles(plhs[1]) + 2 * N * b;
    for (size_t j = 0; j < 2 * N; j++) {
      pe[j] = (float)f[j];
      po[j] = (float)oc[j];
    }
    mxGetDoubles(plhs[3])[b] = s_b;
    if (nlhs > 4)
      ofx_opo_read_average(run, b, mxGetDoubles(plhs[4]) + N * b,
                           nlhs > 5 ? mxGetDoubles(plhs[5]) + N * b : nullptr, nullptr);
  }
  ofx_opo_destroy(run);
}
