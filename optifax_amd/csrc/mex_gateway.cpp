// OPOmexBatch MEX gateway: the drop-in for the reference's
// toolbox/lib/MEXlib/src/OPO_mex_BATCH.cpp:19-141 with the same 14-input /
// 3-output contract that OpticalSim.run calls (OpticalSim.m:237-252):
//
//   [Et, ApFT, stepMods] = OPOmexBatch(Et, alpha_w, G, w0, bdiffw0, h, NSteps,
//                                      dt, beta_op, maxErr, minErr, sel, ApFT, stepMods)
//
// Inputs are host arrays (MATLAB's gpuArray needs CUDA, so on an MI355X host
// OpticalSim keeps single-precision host arrays); they are copied to the HIP
// device, one crystal pass runs on it, and new host arrays are returned.  As
// in the reference (:42-56,120-122) the inputs are not modified.
// Built with MATLAB:  mex -R2018a mex_gateway.cpp -I<repo>/include -L<lib> -lofx_hip
// Built without MATLAB against the in-repo mx shim (-DOFX_MX_SHIM).
#ifdef OFX_MX_SHIM
#include "../../include/optifax_mex.h"
#else
#include "mex.h"
#endif
#include <string.h>

#include <vector>

#include "../../include/optifax_hip.h"

namespace {

// the reference's only check (OPO_mex_BATCH.cpp:31-32,38-40)
const char* kIdInvalid = "parallel:gpu:mexGPUExample:InvalidInput";
const char* kMsgInvalid = "Invalid input to MEX file.";

ofx_context* g_ctx = nullptr;
void release_context() {
  if (g_ctx) ofx_destroy(g_ctx);
  g_ctx = nullptr;
}

std::vector<float> real_values(const mxArray* a, const char* name) {
  if (mxIsComplex(a)) mexErrMsgIdAndTxt("optifax:OPOmexBatch:type", "%s must be real", name);
  const size_t n = mxGetNumberOfElements(a);
  std::vector<float> v(n);
  if (mxGetClassID(a) == mxSINGLE_CLASS) {
    memcpy(v.data(), mxGetSingles(a), n * sizeof(float));
  } else if (mxGetClassID(a) == mxDOUBLE_CLASS) {
    const double* d = mxGetDoubles(a);
    for (size_t i = 0; i < n; i++) v[i] = (float)d[i];
  } else {
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:type", "%s must be single or double", name);
  }
  return v;
}

std::vector<float> complex_values(const mxArray* a, const char* name) {
  const size_t n = mxGetNumberOfElements(a);
  std::vector<float> v(2 * n, 0.0f);
  if (mxGetClassID(a) == mxSINGLE_CLASS && mxIsComplex(a)) {
    memcpy(v.data(), mxGetComplexSingles(a), 2 * n * sizeof(float));
  } else if (mxGetClassID(a) == mxSINGLE_CLASS) {
    const float* f = mxGetSingles(a);
    for (size_t i = 0; i < n; i++) v[2 * i] = f[i];
  } else if (mxGetClassID(a) == mxDOUBLE_CLASS && mxIsComplex(a)) {
    const mxComplexDouble* d = mxGetComplexDoubles(a);
    for (size_t i = 0; i < n; i++) {
      v[2 * i] = (float)d[i].real;
      v[2 * i + 1] = (float)d[i].imag;
    }
  } else if (mxGetClassID(a) == mxDOUBLE_CLASS) {
    const double* d = mxGetDoubles(a);
    for (size_t i = 0; i < n; i++) v[2 * i] = (float)d[i];
  } else {
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:type", "%s must be single or double", name);
  }
  return v;
}

mxArray* new_like(const mxArray* a, mxComplexity cpx) {
  return mxCreateNumericArray(mxGetNumberOfDimensions(a), mxGetDimensions(a), mxSINGLE_CLASS, cpx);
}

}  // namespace

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 14) mexErrMsgIdAndTxt(kIdInvalid, kMsgInvalid);
  (void)nlhs;
  const mxArray* aEt = prhs[0];
  const mxArray* aBeta = prhs[8];
  const size_t N = mxGetNumberOfElements(aBeta);
  const size_t nE = mxGetNumberOfElements(aEt);
  if (N == 0 || nE % N != 0)
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:size", "numel(Et)=%zu is not a multiple of numel(beta_op)=%zu",
                      nE, N);
  const int B = (int)(nE / N);
  if (B == 0) {
    // an empty batch: nothing to solve; the outputs are copies of the
    // inputs, as the reference returns its copied arrays (OPO_mex_BATCH.cpp:42,55-56,120-122)
    plhs[0] = mxDuplicateArray(aEt);
    plhs[1] = mxDuplicateArray(prhs[12]);
    plhs[2] = mxDuplicateArray(prhs[13]);
    return;
  }
  std::vector<float> Et = complex_values(aEt, "Et");
  std::vector<float> alpha = real_values(prhs[1], "alpha_w");
  std::vector<float> G = real_values(prhs[2], "G");
  std::vector<float> beta = real_values(aBeta, "beta_op");
  if (alpha.size() != N)
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:size", "numel(alpha_w) must equal numel(beta_op)");
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
  int g_batch = 0;
  if (G.size() == (size_t)p.nsteps) g_batch = 1;
  else if (G.size() == (size_t)p.nsteps * (size_t)B) g_batch = B;
  else
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:size", "numel(G)=%zu must be NSteps or NSteps*B", G.size());
  const mxArray* aAp = prhs[12];
  const size_t nAp = mxGetNumberOfElements(aAp);
  if (nAp % (N * (size_t)B) != 0)
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:size", "numel(ApFT) must be N*(P-1)*B");
  p.n_cols = (int)(nAp / (N * (size_t)B));
  std::vector<float> Ap = complex_values(aAp, "ApFT");
  const mxArray* aSm = prhs[13];
  std::vector<float> sm1 = real_values(aSm, "stepMods");
  if (sm1.size() != (size_t)p.nsteps)
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:size", "numel(stepMods) must equal NSteps");
  std::vector<float> sm((size_t)p.nsteps * (size_t)B);
  for (int b = 0; b < B; b++) memcpy(&sm[(size_t)b * p.nsteps], sm1.data(), sm1.size() * 4);
  std::vector<int> status((size_t)B, 0);

  if (!g_ctx) {
    int st = ofx_create(0, &g_ctx);
    if (st) mexErrMsgIdAndTxt("optifax:OPOmexBatch:device", "no usable HIP device (code %d)", st);
    mexAtExit(release_context);
  }
  int st = ofx_crystal_pass_host(g_ctx, &p, B, Et.data(), alpha.data(), G.data(), g_batch,
                                 beta.data(), p.n_cols ? Ap.data() : nullptr, sm.data(),
                                 status.data(), nullptr);
  if (st != OFX_OK && st != OFX_E_NONFINITE)
    mexErrMsgIdAndTxt("optifax:OPOmexBatch:solver", "%s", ofx_last_error(g_ctx));

  plhs[0] = new_like(aEt, mxCOMPLEX);
  memcpy(mxGetComplexSingles(plhs[0]), Et.data(), Et.size() * sizeof(float));
  plhs[1] = new_like(aAp, mxCOMPLEX);
  memcpy(mxGetComplexSingles(plhs[1]), Ap.data(), Ap.size() * sizeof(float));
  plhs[2] = new_like(aSm, mxREAL);
  // the reference writes one shared stepMods row for the whole batch
  // (OPO_CUDA_BATCH_XTAL.cu:277-280); instance 0's record is returned
  memcpy(mxGetSingles(plhs[2]), sm.data(), sm1.size() * sizeof(float));
}
