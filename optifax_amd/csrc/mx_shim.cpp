// Subset of the MATLAB C Matrix API (same names and signatures as MATLAB's
// matrix.h / mex.h, R2018a interleaved-complex layout) so the MEX gateways in
// this directory can be built and exercised without MATLAB.  Arrays are
// column-major with interleaved complex data.  mexErrMsgIdAndTxt throws; the
// exception is turned into a return code by ofx_mex_call().
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/optifax_mex.h"

struct mxArray_tag {
  mxClassID cls = mxUNKNOWN_CLASS;
  mxComplexity cpx = mxREAL;
  std::vector<mwSize> dims;
  std::vector<unsigned char> data;
};

namespace {
struct MexError {
  std::string id, msg;
};
std::vector<void (*)(void)> g_exit_hooks;

size_t elem_size(mxClassID c) {
  switch (c) {
    case mxDOUBLE_CLASS:
    case mxINT64_CLASS:
    case mxUINT64_CLASS:
      return 8;
    case mxSINGLE_CLASS:
    case mxINT32_CLASS:
    case mxUINT32_CLASS:
      return 4;
    case mxINT16_CLASS:
    case mxUINT16_CLASS:
      return 2;
    case mxINT8_CLASS:
    case mxUINT8_CLASS:
    case mxLOGICAL_CLASS:
      return 1;
    default:
      return 0;
  }
}
}  // namespace

extern "C" {

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid,
                              mxComplexity flag) {
  const size_t es = elem_size(classid);
  if (!es) return nullptr;
  mxArray* a = new (std::nothrow) mxArray_tag();
  if (!a) return nullptr;
  a->cls = classid;
  a->cpx = flag;
  size_t n = 1;
  for (mwSize i = 0; i < ndim; i++) {
    a->dims.push_back(dims[i]);
    n *= dims[i];
  }
  while (a->dims.size() < 2) a->dims.push_back(1);
  a->data.assign(n * es * (flag == mxCOMPLEX ? 2 : 1), 0);
  return a;
}

mxArray* mxDuplicateArray(const mxArray* in) { return in ? new mxArray_tag(*in) : nullptr; }
void mxDestroyArray(mxArray* pa) { delete pa; }

size_t mxGetNumberOfElements(const mxArray* pa) {
  size_t n = 1;
  for (mwSize d : pa->dims) n *= d;
  return n;
}
mwSize mxGetNumberOfDimensions(const mxArray* pa) { return pa->dims.size(); }
const mwSize* mxGetDimensions(const mxArray* pa) { return pa->dims.data(); }
mxClassID mxGetClassID(const mxArray* pa) { return pa->cls; }
int mxIsComplex(const mxArray* pa) { return pa->cpx == mxCOMPLEX; }
void* mxGetData(const mxArray* pa) { return (void*)pa->data.data(); }

double mxGetScalar(const mxArray* pa) {
  if (!pa || pa->data.empty()) return 0.0;
  const void* p = pa->data.data();
  switch (pa->cls) {
    case mxDOUBLE_CLASS: return *(const double*)p;
    case mxSINGLE_CLASS: return *(const float*)p;
    case mxINT32_CLASS: return *(const int*)p;
    case mxUINT32_CLASS: return *(const unsigned*)p;
    case mxINT64_CLASS: return (double)*(const long long*)p;
    case mxUINT64_CLASS: return (double)*(const unsigned long long*)p;
    case mxINT16_CLASS: return *(const short*)p;
    case mxUINT16_CLASS: return *(const unsigned short*)p;
    case mxINT8_CLASS: return *(const signed char*)p;
    case mxUINT8_CLASS:
    case mxLOGICAL_CLASS: return *(const unsigned char*)p;
    default: return 0.0;
  }
}

float* mxGetSingles(const mxArray* pa) {
  return (pa->cls == mxSINGLE_CLASS && pa->cpx == mxREAL) ? (float*)pa->data.data() : nullptr;
}
double* mxGetDoubles(const mxArray* pa) {
  return (pa->cls == mxDOUBLE_CLASS && pa->cpx == mxREAL) ? (double*)pa->data.data() : nullptr;
}
mxComplexSingle* mxGetComplexSingles(const mxArray* pa) {
  return (pa->cls == mxSINGLE_CLASS && pa->cpx == mxCOMPLEX) ? (mxComplexSingle*)pa->data.data()
                                                             : nullptr;
}
mxComplexDouble* mxGetComplexDoubles(const mxArray* pa) {
  return (pa->cls == mxDOUBLE_CLASS && pa->cpx == mxCOMPLEX) ? (mxComplexDouble*)pa->data.data()
                                                             : nullptr;
}

void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, errormsg);
  vsnprintf(buf, sizeof buf, errormsg, ap);
  va_end(ap);
  throw MexError{errorid ? errorid : "", buf};
}

int mexAtExit(void (*exit_fcn)(void)) {
  if (exit_fcn) g_exit_hooks.push_back(exit_fcn);
  return 0;
}

void ofx_mex_clear(void) {
  for (auto f : g_exit_hooks) f();
  g_exit_hooks.clear();
}

int ofx_mex_call(void (*gateway)(int, mxArray**, int, const mxArray**), int nlhs, mxArray* plhs[],
                 int nrhs, const mxArray* prhs[], char* id, int id_len, char* msg, int msg_len) {
  try {
    gateway(nlhs, plhs, nrhs, prhs);
  } catch (const MexError& e) {
    if (id && id_len > 0) snprintf(id, (size_t)id_len, "%s", e.id.c_str());
    if (msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "%s", e.msg.c_str());
    return -1;
  }
  return 0;
}

}  // extern "C"
