// CPU emulation backend for libofx_hip's kernel body (TEST INFRASTRUCTURE).
//
// Compiles optifax_amd/csrc/ofx_pass_impl.h with OFX_EMU: every workgroup is
// one CPU thread executing the kernel's phases sequentially (blockDim = 1),
// cluster barriers become a condition variable over the same counters, and
// "device memory" is host memory.  Linked with the unchanged runtime layer
// (ofx_runtime.cpp) into tests/emu/libofx_emu.so, so tests can run the exact
// indexing, scheduling and step-control code of the GPU path against the
// oracle on a machine without a GPU.  The product never loads this library.
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/optifax_hip.h"
#include "../../optifax_amd/csrc/ofx_backend.h"
#include "../../optifax_amd/csrc/ofx_pass_impl.h"

namespace ofx {

static std::mutex g_mu;
static std::condition_variable g_cv;

bool emu_cluster_barrier(const ClusterDesc* cd, unsigned target, unsigned long long timeout) {
  std::unique_lock<std::mutex> lk(g_mu);
  (*cd->bar)++;
  g_cv.notify_all();
  // the abort flag is set without the lock: re-check it every millisecond
  const unsigned long long t0 = emu_clock();
  while (!(*cd->bar >= target || __atomic_load_n(cd->abort_flag, __ATOMIC_RELAXED))) {
    if (emu_clock() - t0 > timeout) {
      __atomic_store_n(cd->abort_flag, 1, __ATOMIC_RELAXED);
      break;
    }
    g_cv.wait_for(lk, std::chrono::milliseconds(1));
  }
  return __atomic_load_n(cd->abort_flag, __ATOMIC_RELAXED) == 0;
}

// 100 MHz, the wall-clock rate reported by be_create
unsigned long long emu_clock() {
  return (unsigned long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch()).count() / 10ULL;
}

struct Backend {
  std::string err;
  double last_ms = 0.0;
};

int be_create(int device, Backend** out, DeviceInfo* info) {
  if (device != 0) return OFX_E_ARG;
  *out = new Backend();
  memset(info, 0, sizeof *info);
  strcpy(info->name, "cpu-emulation");
  const char* cus = getenv("OFX_EMU_CUS");
  info->cus = cus ? atoi(cus) : 16;
  info->wallclock_khz = 100000;
  info->lds_max = 160 * 1024;
  info->nxcd = 1;
  return OFX_OK;
}
void be_destroy(Backend* b) { delete b; }
const char* be_error(Backend* b) { return b->err.c_str(); }
int be_malloc(Backend* b, void** p, size_t bytes) {
  *p = calloc(1, bytes);
  if (!*p) { b->err = "calloc failed"; return OFX_E_NOMEM; }
  return OFX_OK;
}
void be_free(Backend*, void* p) { free(p); }
int be_h2d(Backend*, void* dst, const void* src, size_t bytes) { memcpy(dst, src, bytes); return OFX_OK; }
int be_d2h(Backend*, void* dst, const void* src, size_t bytes) { memcpy(dst, src, bytes); return OFX_OK; }
int be_sync(Backend*) { return OFX_OK; }
int be_blocks_per_cu(Backend*, int, int, int, size_t lds) { return lds <= 160 * 1024 ? 1 : 0; }

// the RK23 variant's workgroup body (ofx_backend_emu_rkl.cpp)
void (*emu_rkl_fn(int precision, int xmode))(const LaunchArgs*, int, size_t);

template <typename R, int XM>
static void run_wg(const LaunchArgs* a, int wg, size_t lds) {
  std::vector<unsigned char> mem(lds + 64);
  double red[64];
  int sflag[4];
  cluster_main<R, XM>(*a, wg, reinterpret_cast<cpx<R>*>(mem.data()), red, sflag);
}

int be_launch(Backend* b, const LaunchArgs& a, int precision, int xmode, int grid, int block,
              size_t lds, unsigned* const* zp, const size_t* zw, int nz, unsigned* ep, int nep,
              const void* key, int replays) {
  (void)block;
  (void)key;
  const int reps = replays > 0 ? replays : 1;
  for (int rep = 0; rep < reps; rep++) {
    for (int i = 0; i < nz; i++) memset(zp[i], 0, zw[i] * sizeof(unsigned));
    for (int j = 0; ep && j < nep; j++) {
      ep[2 * j] += ep[2 * j + 1];
      ep[2 * j + 1] = 0u;
    }
    std::vector<std::thread> th;
    const bool bar = (xmode & 1) == XM_BARRIER;
    for (int wg = 0; wg < grid; wg++) {
      if (xmode & XMODE_RK23_LDS)
        th.emplace_back(emu_rkl_fn(precision, xmode & 1), &a, wg, lds);
      else if (precision == OFX_F64)
        th.emplace_back(bar ? run_wg<double, XM_BARRIER> : run_wg<double, XM_TAGS>, &a, wg, lds);
      else
        th.emplace_back(bar ? run_wg<float, XM_BARRIER> : run_wg<float, XM_TAGS>, &a, wg, lds);
    }
    for (auto& t : th) t.join();
  }
  b->last_ms = 0.0;
  return OFX_OK;
}
int be_last_ms(Backend* b, double* ms) { *ms = b->last_ms; return OFX_OK; }
void be_forget_graph(Backend*, const void*) {}

// the no-allocation entry point's array, in host memory here
alignas(256) static unsigned char g_emu_scratch[OFX_NOALLOC_SCRATCH_BYTES];
int be_scratch(Backend*, unsigned char** ptr, size_t* bytes) {
  *ptr = g_emu_scratch;
  *bytes = sizeof g_emu_scratch;
  return OFX_OK;
}
int be_fill_zero(Backend*, void* p, size_t bytes) { memset(p, 0, bytes); return OFX_OK; }
int be_widen(Backend*, double* dst, const float* src, size_t n) {
  for (size_t i = 0; i < n; i++) dst[i] = (double)src[i];
  return OFX_OK;
}
int be_d2d(Backend*, void* dst, const void* src, size_t bytes) { memcpy(dst, src, bytes); return OFX_OK; }

}  // namespace ofx
