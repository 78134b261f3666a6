// HIP backend of libofx_hip: device setup, memory, the persistent pass kernel
// (body in ofx_pass_impl.h), graph-safe scratch clearing, timing and hipGraph
// capture/replay.  gfx950 only.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "../../include/optifax_hip.h"
#include "ofx_backend.h"
#include "ofx_pass_impl.h"

namespace ofx {

template <typename R, int XM>
__global__ void __launch_bounds__(512) pass_kernel(LaunchArgs a) {
  extern __shared__ __align__(16) unsigned char ofx_lds_raw[];
  __shared__ double red[64];
  __shared__ int sflag[4];
  cluster_main<R, XM>(a, (int)blockIdx.x, reinterpret_cast<cpx<R>*>(ofx_lds_raw), red, sflag);
}

// the RK23 variant's kernels (ofx_backend_rkl.hip)
const void* rkl_kernel(int precision, int xmode);

// the kernel instantiation for a precision and xmode (exchange mode in bit 0,
// XMODE_RK23_LDS in bit 1)
static const void* pass_fn(int precision, int xmode) {
  if (xmode & XMODE_RK23_LDS) return rkl_kernel(precision, xmode & 1);
  xmode &= 1;
  if (precision == OFX_F64)
    return xmode == XM_BARRIER ? (const void*)pass_kernel<double, XM_BARRIER>
                               : (const void*)pass_kernel<double, XM_TAGS>;
  return xmode == XM_BARRIER ? (const void*)pass_kernel<float, XM_BARRIER>
                             : (const void*)pass_kernel<float, XM_TAGS>;
}

// zero scratch words with vector stores (graph-safe replacement for memset)
// and advance the clusters' tag epochs past the tags of the previous launch
struct ZeroList {
  unsigned* p[16];
  unsigned w[16];
  int n;
  unsigned* epochs;  // [n_epochs][2] {base, largest exchange count used}
  int n_epochs;
};
__global__ void zero_kernel(ZeroList z) {
  for (int i = 0; i < z.n; i++)
    for (unsigned j = threadIdx.x; j < z.w[i]; j += blockDim.x) z.p[i][j] = 0u;
  for (int j = threadIdx.x; j < z.n_epochs; j += blockDim.x) {
    z.epochs[2 * j] += z.epochs[2 * j + 1];
    z.epochs[2 * j + 1] = 0u;
  }
}

// Device array of the no-allocation entry point (ofx_crystal_pass_noalloc):
// exchange buffers, descriptors and tables of its launches.  Each device
// holds its own copy until the process ends.  Host code never uses this
// name as an address: it is a host-side stand-in for the device symbol, and
// the per-device address from hipGetSymbolAddress (g_scratch) is used
// instead.
static __device__ __align__(256) unsigned char scratch[OFX_NOALLOC_SCRATCH_BYTES];

#define OFX_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char* g_scratch[OFX_MAX_DEVICES];  // device address per device, or null

__global__ void fill_zero_kernel(unsigned* p, size_t words) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ void widen_kernel(double* dst, const float* src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}

struct Backend {
  int device = 0;
  unsigned char* scratch = nullptr;  // this device's address of `scratch`
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  std::string err;
  std::map<const void*, hipGraphExec_t> graphs;
};

static int hip_fail(Backend* b, hipError_t e, const char* what) {
  if (b) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    b->err = buf;
  }
  return OFX_E_HIP;
}

#define HIPCHK(b, call)                                  \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return hip_fail((b), _e, #call); \
  } while (0)

int be_create(int device, Backend** out, DeviceInfo* info) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OFX_E_NODEVICE;
  if (device < 0 || device >= ndev) return OFX_E_ARG;
  Backend* b = new Backend();
  b->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { delete b; return OFX_E_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete b; return OFX_E_HIP; }
  memset(info, 0, sizeof *info);
  snprintf(info->name, sizeof info->name, "%s (%s)", prop.name, prop.gcnArchName);
  info->cus = prop.multiProcessorCount;
  int khz = 0;
  hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device);
  info->wallclock_khz = khz > 0 ? khz : 100000;
  int lmax = 0;
  hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  info->lds_max = lmax > 0 ? lmax : 65536;
  // gfx950 / gfx942 dispatch workgroups round-robin over 8 XCDs (32 CUs each)
  info->nxcd = (strstr(prop.gcnArchName, "gfx95") || strstr(prop.gcnArchName, "gfx94")) &&
                       (info->cus % 8 == 0)
                   ? 8
                   : 1;
  // the stream keeps the default priority
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    delete b;
    return OFX_E_HIP;
  }
  // Raise the dynamic-LDS cap of the pass kernels to the device limit (minus
  // their static shared arrays).  Failures are tolerated: occupancy queries
  // and launch errors report an unusable size later.
  {
    const int dyn = info->lds_max - 1024;
    bool ok = true;
    for (int prec : {OFX_F32, OFX_F64})
      for (int xm : {XM_TAGS, XM_BARRIER, XM_TAGS | XMODE_RK23_LDS, XM_BARRIER | XMODE_RK23_LDS})
        ok = ok && hipFuncSetAttribute(pass_fn(prec, xm), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       dyn) == hipSuccess;
    if (!ok)
      snprintf(info->name + strlen(info->name), sizeof info->name - strlen(info->name),
               " [dyn-LDS attribute not set]");
    (void)hipGetLastError();  // do not leak this status into later launch checks
  }
  snprintf(info->name + strlen(info->name), sizeof info->name - strlen(info->name), " lds=%d",
           info->lds_max);
  // The no-allocation entry point's array: its address on this device, once
  // per device (with the device current), and one clearing of its contents.
  if (device >= OFX_MAX_DEVICES) {
    be_destroy(b);
    return OFX_E_UNSUPPORTED;
  }
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, b->stream, (unsigned*)p,
                           (size_t)OFX_NOALLOC_SCRATCH_BYTES / 4);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) {
        be_destroy(b);
        return OFX_E_HIP;
      }
      g_scratch[device] = (unsigned char*)p;
    }
    b->scratch = g_scratch[device];
  }
  *out = b;
  return OFX_OK;
}

int be_scratch(Backend* b, unsigned char** ptr, size_t* bytes) {
  if (!b || !b->scratch) return OFX_E_UNSUPPORTED;
  *ptr = b->scratch;
  *bytes = (size_t)OFX_NOALLOC_SCRATCH_BYTES;
  return OFX_OK;
}

int be_fill_zero(Backend* b, void* p, size_t bytes) {
  const size_t words = bytes / 4;
  if (!words) return OFX_OK;
  const size_t blocks = std::min<size_t>((words + 255) / 256, 2048);
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, b->stream,
                     (unsigned*)p, words);
  HIPCHK(b, hipGetLastError());
  return OFX_OK;
}

int be_widen(Backend* b, double* dst, const float* src, size_t n) {
  if (!n) return OFX_OK;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)blocks), dim3(256), 0, b->stream, dst, src, n);
  HIPCHK(b, hipGetLastError());
  return OFX_OK;
}

int be_d2d(Backend* b, void* dst, const void* src, size_t bytes) {
  HIPCHK(b, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b->stream));
  return OFX_OK;
}

void be_destroy(Backend* b) {
  if (!b) return;
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
  hipEventDestroy(b->ev0);
  hipEventDestroy(b->ev1);
  hipStreamDestroy(b->stream);
  delete b;
}

const char* be_error(Backend* b) { return b ? b->err.c_str() : "no backend"; }

int be_malloc(Backend* b, void** p, size_t bytes) {
  hipSetDevice(b->device);
  HIPCHK(b, hipMalloc(p, bytes));
  HIPCHK(b, hipMemsetAsync(*p, 0, bytes, b->stream));
  return OFX_OK;
}

void be_free(Backend* b, void* p) {
  if (p) hipFree(p);
  (void)b;
}

int be_h2d(Backend* b, void* dst, const void* src, size_t bytes) {
  HIPCHK(b, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
  return OFX_OK;
}

int be_d2h(Backend* b, void* dst, const void* src, size_t bytes) {
  HIPCHK(b, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, b->stream));
  return OFX_OK;
}

int be_sync(Backend* b) {
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return OFX_OK;
}

int be_blocks_per_cu(Backend* b, int precision, int xmode, int block, size_t lds_bytes) {
  int nb = 0;
  const void* f = pass_fn(precision, xmode);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, lds_bytes) != hipSuccess) return 0;
  (void)b;
  return nb;
}

static int enqueue(Backend* b, const LaunchArgs& a, int precision, int xmode, int grid, int block,
                   size_t lds, unsigned* const* zp, const size_t* zw, int nz, unsigned* ep,
                   int nep) {
  (void)hipGetLastError();
  for (int base = 0; base < nz || base == 0; base += 16) {
    ZeroList z;
    memset(&z, 0, sizeof z);
    z.n = nz - base < 16 ? nz - base : 16;
    if (z.n < 0) z.n = 0;
    for (int i = 0; i < z.n; i++) {
      z.p[i] = zp[base + i];
      z.w[i] = (unsigned)zw[base + i];
    }
    if (base == 0) {
      z.epochs = ep;
      z.n_epochs = ep ? nep : 0;
    }
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, b->stream, z);
    HIPCHK(b, hipGetLastError());
  }
  void* args[] = {(void*)&a};
  HIPCHK(b, hipLaunchKernel(pass_fn(precision, xmode), dim3(grid), dim3(block), args, lds,
                            b->stream));
  return OFX_OK;
}

int be_launch(Backend* b, const LaunchArgs& a, int precision, int xmode, int grid, int block,
              size_t lds, unsigned* const* zp, const size_t* zw, int nz, unsigned* ep, int nep,
              const void* key, int replays) {
  hipSetDevice(b->device);
  if (replays <= 0) {
    HIPCHK(b, hipEventRecord(b->ev0, b->stream));
    int st = enqueue(b, a, precision, xmode, grid, block, lds, zp, zw, nz, ep, nep);
    if (st) return st;
    HIPCHK(b, hipEventRecord(b->ev1, b->stream));
    b->timed = true;
    return OFX_OK;
  }
  hipGraphExec_t exec = nullptr;
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    hipGraph_t g = nullptr;
    HIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue(b, a, precision, xmode, grid, block, lds, zp, zw, nz, ep, nep);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) return st;
    if (e != hipSuccess) return hip_fail(b, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(b, e, "hipGraphInstantiate");
    b->graphs[key] = exec;
  } else {
    exec = it->second;
  }
  HIPCHK(b, hipEventRecord(b->ev0, b->stream));
  for (int i = 0; i < replays; i++) HIPCHK(b, hipGraphLaunch(exec, b->stream));
  HIPCHK(b, hipEventRecord(b->ev1, b->stream));
  b->timed = true;
  return OFX_OK;
}

int be_last_ms(Backend* b, double* ms) {
  if (!b->timed) return OFX_E_ARG;
  float f = 0.f;
  HIPCHK(b, hipEventSynchronize(b->ev1));
  HIPCHK(b, hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return OFX_OK;
}

void be_forget_graph(Backend* b, const void* key) {
  auto it = b->graphs.find(key);
  if (it != b->graphs.end()) {
    hipStreamSynchronize(b->stream);
    hipGraphExecDestroy(it->second);
    b->graphs.erase(it);
  }
}

}  // namespace ofx
