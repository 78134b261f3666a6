/*
 * optifax_hip.h -- C ABI of libofx_hip, the MI355X (gfx950) implementation of
 * OptiFax's adaptive split-step chi(2) crystal solver and OPO round-trip loop.
 *
 * Plain C types only.  Two layers:
 *
 *  1. ofx_crystal_pass_host(): one crystal pass over a batch of fields held
 *     in host memory.  It is the computation behind the reference's MEX
 *     entry point
 *        [Et,ApFT,stepMods] = OPOmexBatch(Et,alpha_w,G,w0,bdiffw0,h,NSteps,dt,
 *                                         beta_op,maxErr,minErr,sel,ApFT,stepMods)
 *     (toolbox/lib/MEXlib/src/OPO_mex_BATCH.cpp:19-141, which calls
 *      OPO_TEST_CUDA, OPO_CUDA_BATCH_XTAL.cu:373-525, declared in
 *      OPO_BATCH_HEADER.hpp:8-30).  The MEX gateway in
 *     optifax_amd/csrc/mex_gateway.cpp (include/optifax_mex.h) is a thin
 *     wrapper over it.
 *
 *  2. ofx_opo_*: a device-resident batch of OPO instances that runs whole
 *     round trips (crystal pass + output coupling + cavity detuning + pump)
 *     inside one persistent kernel.  It replaces the MATLAB loop
 *     OpticalSim.run (toolbox/lib/classes/OpticalSim.m:207-301) plus its per
 *     trip solver call (:237-252); there is no C counterpart in the reference.
 *
 * Fields are interleaved complex (re, im) arrays of length 2*N, time samples
 * in fftshifted order (t = 0 at index 0), exactly as OpticalSim passes EtShift
 * (OpticalSim.m:234).  Spectral tables are in FFT order (OpticalSim.m:223).
 *
 * All functions return OFX_OK (0) or a negative OFX_E_* code; the message of
 * the last failure is available from ofx_last_error().
 */
#ifndef OPTIFAX_HIP_H
#define OPTIFAX_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define OFX_OK 0
#define OFX_E_ARG (-1)          /* invalid argument (shape, size, value)          */
#define OFX_E_HIP (-2)          /* HIP runtime error                              */
#define OFX_E_NOMEM (-3)        /* device or host allocation failed               */
#define OFX_E_NONFINITE (-4)    /* an instance produced NaN/Inf (see statuses)    */
#define OFX_E_TIMEOUT (-5)      /* a cluster barrier timed out (see statuses)     */
#define OFX_E_UNSUPPORTED (-6)  /* option not implemented on this device          */
#define OFX_E_NODEVICE (-7)     /* no HIP device                                  */

/* per-instance status after a run */
#define OFX_STATUS_RUNNING 0
#define OFX_STATUS_DONE 1
#define OFX_STATUS_CONVERGED 2
#define OFX_STATUS_NONFINITE 3
#define OFX_STATUS_TIMEOUT 4

#define OFX_F32 0 /* complex64 fields, FP64 phase arguments and norms  */
#define OFX_F64 1 /* complex128 fields throughout                       */

#define OFX_METHOD_HEUN 0 /* adaptive Heun, NEE_CPU.m:39-94 semantics         */
#define OFX_METHOD_RK23 1 /* Bogacki-Shampine, ode23/OPO_CUDA_ODE23.cu:375-557 */

typedef struct ofx_context ofx_context;
typedef struct ofx_opo ofx_opo;

/* Scalars of one crystal pass: arguments 4-8 and 10-12 of OPOmexBatch
 * (OPO_mex_BATCH.cpp:45-52) plus the derived sizes (:71-72). */
typedef struct {
  int n;          /* points N = numel(beta_op), power of two, 64..65536  */
  int nsteps;     /* NSteps = floor(L/h)                                  */
  double w0;      /* reference angular frequency [rad/s]                  */
  double bdiffw0; /* beta0 - beta1*w0 [1/m]                               */
  double h;       /* minimal step [m]                                     */
  double dt;      /* time step [s]                                        */
  double max_err; /* reject threshold, percent                            */
  double min_err; /* grow threshold, percent                              */
  double sel;     /* minimal steps per ApFT chunk, > 1                    */
  int n_cols;     /* ApFT columns (ProgressPlots-1); 0 = do not write     */
  int method;     /* OFX_METHOD_*                                         */
} ofx_pass_params;

/* Optional launch geometry.  Zero fields mean "autotuned/default". */
typedef struct {
  int cluster;    /* workgroups per instance (power of two)  */
  int n1;         /* four-step split N = n1*n2 (power of two) */
  int block;      /* threads per workgroup (64..512)          */
  int exchange;   /* how a workgroup learns that its cluster's exchange data
                     are complete: 1 poll the data's own tags, 2 a counter
                     barrier first; 0 auto (tags from 64 workgroups up) */
} ofx_tuning;

int ofx_create(int device, ofx_context** out);
void ofx_destroy(ofx_context* ctx);
const char* ofx_last_error(const ofx_context* ctx);
/* device name, compute units, wall-clock rate (kHz) */
int ofx_device_info(const ofx_context* ctx, char* name, int name_len, int* cus, int* wallclock_khz);

/*
 * One crystal pass over `batch` fields in host memory (OPOmexBatch semantics,
 * Heun step control of NEE_CPU.m:39-94 applied independently per instance;
 * batch 0 returns OFX_OK with nothing done, as an empty Et does there).
 *  Et       [2*n*batch] in/out, complex64 interleaved, fftshifted time
 *  alpha_w  [n]          T^(1/NSteps) in FFT order (NonlinearCrystal.m:141-144)
 *  G        [nsteps*g_batch] coupling per minimal step; g_batch 1 (shared) or batch
 *  beta_op  [n]          phi_rel/L*h in FFT order (OpticalSim.m:223)
 *  ApFT     [2*n*n_cols*batch] or NULL: spectrum at each chunk end,
 *           instance b column c at offset 2*n*(b*n_cols + c)
 *  stepMods [nsteps*batch] or NULL: stepmod accepted at each 1-based k (k-1)
 *  status   [batch] or NULL: OFX_STATUS_* per instance
 */
int ofx_crystal_pass_host(ofx_context* ctx, const ofx_pass_params* p, int batch, float* Et,
                          const float* alpha_w, const float* G, int g_batch,
                          const float* beta_op, float* ApFT, float* stepMods, int* status,
                          const ofx_tuning* tune);

/*
 * The same pass with no allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.
 *
 * Every array is a DEVICE pointer (float complex64 fields and float tables,
 * laid out as for ofx_crystal_pass_host) and the pass is computed in place:
 * Et, ApFT and stepMods are overwritten where the pass writes them, status
 * (or NULL) receives OFX_STATUS_* per instance.  The inputs must be complete
 * on the device before the call; the call runs on the context's stream and
 * returns when the pass is done, with the same return codes as
 * ofx_crystal_pass_host.  Default launch geometry (no tuning argument).
 *
 * Exchange buffers, descriptors and tables live in one device array of
 * OFX_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs
 * on until the process ends; the call allocates nothing (device or host).
 * A configuration whose buffers do not fit the array (roughly N*batch up to
 * 2^17 with Heun, less with RK23 or long crystals), or batch > 16, is
 * rejected with OFX_E_ARG and a message.  Calls in one process are
 * serialized by the library (they share the array, which keeps the last
 * call's contents; every call clears what it needs cleared).
 */
#define OFX_NOALLOC_SCRATCH_BYTES (16u << 20)
int ofx_crystal_pass_noalloc(ofx_context* ctx, const ofx_pass_params* p, int batch, float* Et,
                             const float* alpha_w, const float* G, int g_batch,
                             const float* beta_op, float* ApFT, float* stepMods, int* status);

/* Same, FP64 fields (complex128) and FP64 tables; used for exact parity. */
int ofx_crystal_pass_host_f64(ofx_context* ctx, const ofx_pass_params* p, int batch, double* Et,
                              const double* alpha_w, const double* G, int g_batch,
                              const double* beta_op, double* ApFT, double* stepMods,
                              int* status, const ofx_tuning* tune);

/* ---------------- device-resident OPO batch ---------------- */

/* One OPO instance.  All host arrays are FP64 and are copied (converted to
 * the batch precision) at ofx_opo_create; FFT order for spectra. */
typedef struct {
  ofx_pass_params pass;
  double delay;          /* cavity detuning per trip [s] (applyGD, OpticalPulse.m:256-261) */
  const double* beta;    /* [n]   phi_rel/L*h                                  */
  const double* alpha;   /* [n]   T^(1/NSteps)                                 */
  const double* G;       /* [nsteps]                                           */
  const double* field;   /* [2n]  crystal input of the first trip (fftshifted) */
  const double* pump;    /* [2n]  pump spectrum in the crystal, or NULL        */
  const double* cav;     /* [2n]  crystal-out -> next crystal-in transfer, or NULL (single passes) */
  const double* ocm;     /* [2n]  crystal-out -> output-coupled spectrum, or NULL */
  int cluster;           /* this instance's workgroups (power of two), 0 = batch tuning / automatic */
  int n1;                /* this instance's four-step split, 0 = batch tuning / automatic */
} ofx_opo_spec;

int ofx_opo_create(ofx_context* ctx, int precision, int n_inst, const ofx_opo_spec* specs,
                   int trip_cap, const ofx_tuning* tune, ofx_opo** out);
void ofx_opo_destroy(ofx_opo* run);
/* enqueue `trips` round trips (0 = one plain crystal pass) on the context
 * stream; instances stop early once the trip-to-trip relative L2 change of
 * the output-coupled power spectrum is below conv_tol (<= 0: never).  With
 * conv_tol > 0, an instance that converged in an earlier launch is skipped
 * until ofx_opo_write_field gives it a new field. */
int ofx_opo_launch(ofx_opo* run, int trips, double conv_tol);
int ofx_opo_sync(ofx_opo* run);
/* capture `trips` round trips as a hipGraph (once) and replay it `replays` times */
int ofx_opo_graph_replay(ofx_opo* run, int trips, double conv_tol, int replays);
/* time of the last ofx_opo_launch / replay span on the stream, milliseconds */
int ofx_opo_last_ms(ofx_opo* run, double* ms);
/* read back one instance (any pointer may be NULL).
 *  field [2n], oc_spec [2n], trip_log [4*trip_cap] (conv, oc energy, attempts, accepted),
 *  apft [2n*n_cols], stepmods [nsteps] (after round trips: the last trip's
 *  row, zero where no step ended, as StepSizeModifiers(trip,:),
 *  OpticalSim.m:172,237), counters [2] (attempts, accepted, summed over the
 *  run's launches) */
int ofx_opo_read(ofx_opo* run, int inst, double* field, double* oc_spec, double* trip_log,
                 int* status, int* trips_done, long long* counters, double* apft,
                 double* stepmods);
/* Run averages of one cavity instance, summed on the device in FP64 as each
 * trip ends (either pointer may be NULL): oc_esd [n] = mean of the
 * output-coupled |E_oc(w_j)|^2 and xout_esd [n] = mean of the crystal-output
 * |E_x(w_j)|^2, FFT order, over the trips since creation or the last
 * ofx_opo_reset_average; *trips = the trips averaged (all zeros for 0).
 * Replaces the ESDOutAverage, ESDPumpDepAverage and ESDIdlerICAverage sums of
 * OpticalSim.averageSimPulses (OpticalSim.m:302-313; the last two are the
 * crystal-output ESD behind a wavelength mask, applied by the caller), without
 * their per-trip gather to the host. */
int ofx_opo_read_average(ofx_opo* run, int inst, double* oc_esd, double* xout_esd, int* trips);
/* clear both averages (the next trip starts new ones) */
int ofx_opo_reset_average(ofx_opo* run, int inst);
/* Keep the crystal-output spectrum (FFT order, inside the crystal) and the
 * step multipliers of the last `trips` round trips of every cavity instance
 * in device memory, from the next trip on (trips x (n complex + nsteps
 * floats) per instance; once per run).  Replaces StoredPulses
 * (OpticalSim.m:173-180,268-272) and the per-trip rows of StepSizeModifiers
 * (OpticalSim.m:172,237), which gather every trip to the host. */
int ofx_opo_keep_history(ofx_opo* run, int trips);
/* cumulative 0-based trip `trip` of one instance, if held (either pointer
 * may be NULL): spectrum [2n]; stepmods [nsteps], the multiplier accepted at
 * each 1-based minimal step k in entry k-1 and 0 elsewhere, as the row
 * StepSizeModifiers(trip, :) */
int ofx_opo_read_history(ofx_opo* run, int inst, int trip, double* spectrum, double* stepmods);
/* overwrite the crystal-input field of one instance (warm start / checkpoint
 * restore); resets its status to running */
int ofx_opo_write_field(ofx_opo* run, int inst, const double* field);
/* geometry actually used: cluster size, n1, block, grid, LDS bytes */
int ofx_opo_geometry(const ofx_opo* run, int inst, int* cluster, int* n1, int* block, int* grid,
                     int* lds_bytes);
/* hardware XCC id of every launched workgroup of the last launch (diagnostic) */
int ofx_opo_xcc_ids(ofx_opo* run, int* ids, int cap, int* count);
/* longest time one workgroup waits for its cluster's exchange data before
 * the cluster gives up (default 10 s); a launch that gives up leaves the
 * instance's field unchanged and ofx_opo_sync returns OFX_E_TIMEOUT.  No
 * reference counterpart (the reference's solver has no cross-block waits). */
int ofx_opo_set_wait_limit(ofx_opo* run, double seconds);

/* ---- spectral surrogate of one crystal pass (MLP, BF16 MFMA / FP32) ------
 * No reference counterpart: the reference has no learned component.  Network
 * n_in -> width -> width -> width -> n_out, GELU after the first three layers,
 * weights [out][in] (PyTorch Linear order).  Public parameter vector, FP32:
 *   W0 [width][n_in], b0 [width], W1 [width][width], b1, W2, b2,
 *   W3 [n_out][width], b3 [n_out].
 * Loss: mean over the batch of ||y - t||_2 / ||t||_2.  Optimiser: Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8).  width and n_out: multiples of 64;
 * batch: a multiple of 64 up to max_batch.  Calls on one object are
 * synchronous and must not overlap. */
typedef struct ofx_mlp ofx_mlp;
int ofx_mlp_create(int device, int n_in, int width, int n_out, int max_batch, ofx_mlp** out);
void ofx_mlp_destroy(ofx_mlp* m);
const char* ofx_mlp_last_error(const ofx_mlp* m);
long long ofx_mlp_num_params(const ofx_mlp* m);
/* load parameters (public layout) and reset the Adam state */
int ofx_mlp_set_params(ofx_mlp* m, const float* params);
int ofx_mlp_get_params(ofx_mlp* m, float* params);
/* inference: x [batch][n_in] -> y [batch][n_out] (host arrays) */
int ofx_mlp_forward(ofx_mlp* m, int batch, const float* x, float* y);
/* loss and its gradient w.r.t. the parameters (public layout); no update */
int ofx_mlp_gradients(ofx_mlp* m, int batch, const float* x, const float* t, float* loss,
                      float* grads);
/* one Adam step on the given batch; *loss = batch loss before the update */
int ofx_mlp_step(ofx_mlp* m, int batch, const float* x, const float* t, float lr, float* loss);
/* keep a training set on the device: X [n][n_in], T [n][n_out] */
int ofx_mlp_upload_dataset(ofx_mlp* m, int n, const float* X, const float* T);
/* `steps` Adam steps; step s uses dataset rows idx[s*batch .. s*batch+batch-1];
 * losses [steps] (may be NULL) */
int ofx_mlp_train(ofx_mlp* m, int batch, int steps, const int* idx, float lr, float* losses);
/* device time of the last ofx_mlp_forward / ofx_mlp_train, milliseconds */
double ofx_mlp_last_ms(const ofx_mlp* m);

/* ---- QPM gain maps (phase-matching analysis) ------------------------------
 * Synchronous calls on `device`; host arrays in and out; *ms (may be NULL)
 * receives the kernel's device time in milliseconds.
 *
 * ofx_qpm_zsum: the sum over the crystal of qpmgain.m:140-149
 *     (dAs_dz = g_coeff .* exp(1i*dk.*z) .* P;  As = sum(dAs_dz.*dz, 3)),
 * without the prefactor g_coeff:
 *     out[o] = sum_r wgt[r] sum_{q < len[r]} exp(i dk[o] (z0[r] + q dz)),
 * o < n_out (complex, interleaved).  A run r stands for len[r] consecutive
 * samples of spacing dz with one weight P(z) dz; runs of one sample give the
 * literal sum over any sampling.  FP64. */
int ofx_qpm_zsum(int device, int n_out, const double* dk, int n_run, const double* z0,
                 const int* len, const double* wgt, double dz, double* out, double* ms);
/* ofx_sinc_gain: the narrowband mismatch factor of sincgain.m (as
 * sincgain_sparse.m:48-70 forms it, in consistent units):
 *     gain[g * n_out + o] = sum_{j < n_m} (sinc((dk0[o] - m[j] kc[g]) length / 2) / m[j])^2,
 * sinc(x) = sin(pi x) / (pi x) (MATLAB's), dk0 and kc in rad/m, length in m.
 * FP64 with the argument reduced mod 2 pi. */
int ofx_sinc_gain(int device, int n_out, const double* dk0, int n_g, const double* kc, int n_m,
                  const double* m, double length, double* gain, double* ms);

/* ---- fibre GNLSE (gnlsefib.m) ------------------------------------------
 * ofx_gnlse_run replaces gnlsefib.m's integration (gnlsefib.m:49-76: the
 * interaction-picture GNLSE right-hand side under ode45 with RelTol / AbsTol
 * and NormControl on, solutions at the requested points).  The host forms
 * the operators as gnlsefib.m:22-45 does; the device integrates.
 * n_inst independent fibres of n points (a power of two, n <= 2^13 with
 * OFX_F32 fields, 2^12 with OFX_F64), one workgroup each, synchronous on
 * `device`.  Arrays (host, FP64; complex interleaved; FFT order):
 *   zs [nsaves >= 3]   requested points, strictly monotone, zs[0] = start
 *   y0 [n_inst][n]     AW at zs[0] (ifft of the temporal field)
 *   L  [n_inst][n]     linear operator i B - alpha / 2 (its real part the
 *                      same at every frequency, else OFX_E_ARG)
 *   W  [n_inst][n]     real: V + w0 with self-steepening, else 1
 *   RW [n_inst][n]     n ifft(fftshift(RT)), or NULL: no Raman term
 *   gamma, fr, dT [n_inst]   (gamma already divided by w0 with self-steepening)
 *   out [n_inst][nsaves][n]  AW at each requested point (interaction picture)
 *   stats [n_inst][4]  accepted steps, rejected steps, evaluations, status
 *                      (0 ok, 1 step below 16 eps(z), 2 max_steps reached,
 *                      3 non-finite)
 * *ms (may be NULL) receives the kernel's device time. */
int ofx_gnlse_run(int device, int precision, int n_inst, int n, int nsaves, const double* zs,
                  const double* y0, const double* L, const double* W, const double* RW,
                  const double* gamma, const double* fr, const double* dT, double rtol,
                  double atol, long long max_steps, double* out, long long* stats, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* OPTIFAX_HIP_H */
