/*
 * optifax_mex.h -- MEX boundary of libofx_hip.
 *
 * The drop-in for the reference's solver handle is the MATLAB MEX entry point
 *
 *   void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
 *
 * built as OPOmexBatch.<mexext> from optifax_amd/csrc/mex_gateway.cpp.  It
 * replaces toolbox/lib/MEXlib/src/OPO_mex_BATCH.cpp:19-141 (built by
 * toolbox/scripts/cuda_compile.m:9-17) with the same 14-input / 3-output
 * contract that OpticalSim.run calls (toolbox/lib/classes/OpticalSim.m:237-252):
 *
 *   [Et, ApFT, stepMods] = OPOmexBatch(Et, alpha_w, G, w0, bdiffw0, h, NSteps,
 *                                      dt, beta_op, maxErr, minErr, sel,
 *                                      ApFT, stepMods)
 *
 * and a second entry point built as OPOmexCavity.<mexext> from
 * optifax_amd/csrc/mex_cavity.cpp (no reference counterpart: it runs whole
 * round trips on the device, replacing the OpticalSim.run loop).
 *
 * With MATLAB, the gateways include MATLAB's own mex.h (R2018a interleaved
 * complex API).  Without MATLAB they are built against the declarations
 * below, implemented by optifax_amd/csrc/mx_shim.cpp: a subset of the
 * documented MATLAB C Matrix API with the same names and signatures, plus
 * ofx_mex_call(), which runs a gateway and turns mexErrMsgIdAndTxt into a
 * return code.
 */
#ifndef OPTIFAX_MEX_H
#define OPTIFAX_MEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;

typedef enum {
  mxUNKNOWN_CLASS = 0,
  mxCELL_CLASS,
  mxSTRUCT_CLASS,
  mxLOGICAL_CLASS,
  mxCHAR_CLASS,
  mxVOID_CLASS,
  mxDOUBLE_CLASS,
  mxSINGLE_CLASS,
  mxINT8_CLASS,
  mxUINT8_CLASS,
  mxINT16_CLASS,
  mxUINT16_CLASS,
  mxINT32_CLASS,
  mxUINT32_CLASS,
  mxINT64_CLASS,
  mxUINT64_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

typedef struct {
  float real, imag;
} mxComplexSingle;
typedef struct {
  double real, imag;
} mxComplexDouble;

/* MATLAB C Matrix API subset (same signatures as MATLAB's matrix.h / mex.h) */
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid,
                              mxComplexity flag);
mxArray* mxDuplicateArray(const mxArray* in);
void mxDestroyArray(mxArray* pa);
size_t mxGetNumberOfElements(const mxArray* pa);
mwSize mxGetNumberOfDimensions(const mxArray* pa);
const mwSize* mxGetDimensions(const mxArray* pa);
mxClassID mxGetClassID(const mxArray* pa);
int mxIsComplex(const mxArray* pa);
double mxGetScalar(const mxArray* pa);
void* mxGetData(const mxArray* pa);
float* mxGetSingles(const mxArray* pa);
double* mxGetDoubles(const mxArray* pa);
mxComplexSingle* mxGetComplexSingles(const mxArray* pa);
mxComplexDouble* mxGetComplexDoubles(const mxArray* pa);
void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
int mexAtExit(void (*exit_fcn)(void));

/* the gateway (defined in each MEX source) */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

/* Shim only: call a gateway; returns 0, or -1 with the error id and message
 * raised through mexErrMsgIdAndTxt copied into id / msg. */
int ofx_mex_call(void (*gateway)(int, mxArray**, int, const mxArray**), int nlhs, mxArray* plhs[],
                 int nrhs, const mxArray* prhs[], char* id, int id_len, char* msg, int msg_len);
/* Shim only: run the exit hooks registered with mexAtExit (MATLAB's clear mex). */
void ofx_mex_clear(void);

#ifdef __cplusplus
}
#endif
#endif /* OPTIFAX_MEX_H */
