// Device backend interface used by the HIP-free runtime layer (ofx_runtime.cpp).
// The product implementation is ofx_backend_hip.hip; tests/emu provides a CPU
// implementation that runs the same kernel body for indexing/control checks.
#pragma once
#include <stddef.h>
#include "ofx_internal.h"

// Bit 1 of a launch's xmode (bit 0 is the exchange mode): the kernel variant
// whose RK23 pass reads the cluster descriptor from a copy in LDS, used when
// every instance of the launch is RK23 (ofx_backend_rkl.hip).
enum { XMODE_RK23_LDS = 2 };

namespace ofx {

struct DeviceInfo {
  char name[128];
  int cus;            // compute units
  int wallclock_khz;  // constant-rate clock used by the barrier timeout
  int lds_max;        // max dynamic LDS bytes per workgroup
  int nxcd;           // accelerator complex dies (workgroups round-robin over them)
};

struct Backend;

int be_create(int device, Backend** out, DeviceInfo* info);
void be_destroy(Backend* b);
const char* be_error(Backend* b);
// zero-filled device memory (stream-ordered)
int be_malloc(Backend* b, void** p, size_t bytes);
void be_free(Backend* b, void* p);
// stream-ordered copies (the host buffer must stay valid until be_sync)
int be_h2d(Backend* b, void* dst, const void* src, size_t bytes);
int be_d2h(Backend* b, void* dst, const void* src, size_t bytes);
int be_sync(Backend* b);
// workgroups of the pass kernel that fit on one CU (0 if the config cannot launch)
int be_blocks_per_cu(Backend* b, int precision, int xmode, int block, size_t lds_bytes);
// Enqueue: zero `nzero` words at `zero_ptrs[i]` and advance the `n_epochs`
// tag epochs at `epochs` (pairs {base, used}: base += used, used = 0) in one
// graph-safe fill kernel, then the pass kernel.  If replays > 0 the pair is captured once into a graph keyed by
// `key` and replayed `replays` times.  Elapsed time of the enqueued span is
// available from be_last_ms after be_sync.
// xmode selects the kernel's exchange mode (0 tagged polling, 1 counter barrier).
int be_launch(Backend* b, const LaunchArgs& a, int precision, int xmode, int grid, int block,
              size_t lds, unsigned* const* zero_ptrs, const size_t* zero_words, int nzero,
              unsigned* epochs, int n_epochs, const void* key, int replays);
int be_last_ms(Backend* b, double* ms);
void be_forget_graph(Backend* b, const void* key);
// The library's fixed device array for the no-allocation entry point: its
// device address on this backend's device (looked up once per device when
// the first backend of that device is created) and its size in bytes.
int be_scratch(Backend* b, unsigned char** ptr, size_t* bytes);
// stream-ordered: zero `bytes` (a multiple of 4) at device address p with a
// fill kernel; dst[i] = (double)src[i] for device arrays; device-to-device copy
int be_fill_zero(Backend* b, void* p, size_t bytes);
int be_widen(Backend* b, double* dst, const float* src, size_t n);
int be_d2d(Backend* b, void* dst, const void* src, size_t bytes);

}  // namespace ofx
