// Internal types shared by the HIP kernels and the host runtime of libofx_hip.
// Not part of the public C ABI (include/optifax_hip.h is).
#pragma once
#ifndef OFX_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace ofx {

// Method codes (match OFX_METHOD_* in optifax_hip.h).
enum { METHOD_HEUN = 0, METHOD_RK23 = 1 };

// Instance status codes written by the kernel (match OFX_STATUS_* values).
enum {
  ST_RUNNING = 0,
  ST_DONE = 1,
  ST_CONVERGED = 2,
  ST_NONFINITE = 3,
  ST_BARRIER_TIMEOUT = 4,
  ST_BAD_ARG = 5,
};

// Everything one instance needs on the device.  One array entry per instance,
// uploaded once per run; pointers are device pointers.
struct InstDesc {
  // grid: N = n1 * n2 points, time index j = c + n1 * r (c column, r row)
  int n, n1, n2, lg1, lg2;
  int nsteps;      // NSteps (number of minimal steps h)
  int n_cols;      // ApFT columns written (0 = none)
  int method;      // METHOD_HEUN or METHOD_RK23
  double w0;       // reference angular frequency [rad/s]
  double bdiffw0;  // beta0 - beta1*w0 [1/m]
  double h;        // minimal step [m]
  double dt;       // time step [s]
  double max_err;  // percent
  double min_err;  // percent
  double sel;      // steps per ApFT chunk
  double delay;    // cavity detuning [s] (round-trip mode)
  // read-only tables (FFT order)
  const double* beta;   // N: per-minimal-step phase phi_rel/L*h
  const double* alpha;  // N: T^(1/NSteps)
  const double* G;      // nsteps: poling-signed coupling
  const void* wt;       // N complex: exp(i*w0*t_j), phase reduced in FP64 (fftshifted time)
  const void* absorb;   // N real: absorbing-boundary factor of t_j (1 outside the zone)
  // in/out
  void* field;          // N complex: fftshifted time field (crystal input / output)
  void* apft;           // n_cols*N complex or null
  float* stepmods;      // nsteps or null
  // round-trip mode (all null in single-pass mode)
  const void* pump;     // N complex: pump spectrum in the crystal (FFT order)
  const void* cav;      // N complex: crystal output -> next crystal input transfer
  const void* ocm;      // N complex: crystal output -> output-coupled spectrum
  void* oc_spec;        // N complex: last output-coupled spectrum (in/out)
  double* esd_sum;      // N: sum over trips of |output-coupled spectrum|^2 (FFT order)
  double* xesd_sum;     // N: sum over trips of |crystal-output spectrum|^2 (FFT order)
  void* hist;           // [hist_cap][N] complex: crystal-output spectra of recent trips, or null
  float* sm_hist;       // [hist_cap][nsteps]: step multipliers of those trips (0 = not accepted)
  int hist_cap;         // trips held; trip t (cumulative, 0-based) sits in slot t % hist_cap
  int pad2;
  double* trip_log;     // [trip_cap][4]: conv metric, oc energy, attempts, accepted
  int trip_cap;
  int pad1;
  int* status;          // ST_* per instance
  int* trips_done;      // per instance, cumulative (also the trip-log base)
  int* fresh;           // per instance: 1 after a field rewrite (no previous spectrum yet)
  long long* counters;  // [0] attempts, [1] accepted (cumulative over the launch)
};

// One cluster = C workgroups cooperating on one instance at a time.
struct ClusterDesc {
  int C;           // workgroups in this cluster
  int first_inst;  // offset into the instance-order list
  int n_inst;      // instances processed in sequence
  int halo;        // halo columns each side
  // Exchange buffers hold tagged elements: every 32-bit word of a complex
  // sample as one 64-bit word (word | tag << 32), two parity halves each.
  void* t1;          // [2][N] tagged complex: forward transpose, N of the cluster
  void* t2;          // [2][N] tagged complex: inverse transpose
  void* slots;       // [2][C][8 waves][4] tagged complex-sized records (doubles)
  unsigned* bar;     // counter barrier (zeroed before every launch)
  int* abort_flag;   // set when a member gives up waiting
  int* xcc;          // C ints: hardware XCC id of each member (placement diagnostic)
  void* kx;          // RK23 only: [2][C][2 sides][3 columns][n2] tagged complex, K4 boundary
  unsigned* epoch;   // [0] tag base of this launch, [1] largest exchange count used by it
};

struct LaunchArgs {
  const InstDesc* inst;
  const ClusterDesc* clusters;
  const int* inst_order;   // instance indices, grouped per cluster
  const int* wg_cluster;   // per workgroup: cluster index or -1 (idle slot)
  const int* wg_rank;      // per workgroup: rank within its cluster
  const void* twiddle;     // NMAX complex: exp(-2 pi i m / NMAX)
  int nmax;                // twiddle table size (largest N in the launch)
  int trips;               // 0: single pass (MEX mode); >0: round trips
  double conv_tol;         // stop an instance when conv metric < tol (<=0 never)
  unsigned long long timeout_ticks;  // barrier spin limit in wall-clock ticks
  int* xcc_ids;            // optional: per workgroup hardware XCC id (diagnostic)
};

}  // namespace ofx
