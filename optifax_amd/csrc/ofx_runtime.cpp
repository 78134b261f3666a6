// C ABI of libofx_hip (include/optifax_hip.h) over the device backend.
// Host-side planning (launch geometry, cluster schedule, uploads) lives here;
// the kernel body is ofx_pass_impl.h and the HIP specifics ofx_backend_hip.hip.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/optifax_hip.h"
#include "ofx_backend.h"
#include "ofx_internal.h"

using namespace ofx;

// What a cached single-pass run was created for: a later OPOmexBatch call
// with the same key reuses its device buffers and only re-uploads the data.
struct PassKey {
  int prec, batch, g_batch, has_tune, apft, stepmods;
  ofx_pass_params p;
  ofx_tuning t;
};

struct ofx_context {
  Backend* be = nullptr;
  DeviceInfo info{};
  std::string err;
  std::map<std::pair<int, int>, void*> tw;  // (nmax, precision) -> table
  ofx_opo* pass_run = nullptr;              // the last single-pass run, kept for reuse
  PassKey pass_key;
};

struct InstBufs {
  void* wt = nullptr;
  void* absorb = nullptr;
  void* field = nullptr;
  void* beta = nullptr;
  void* alpha = nullptr;
  void* G = nullptr;
  void* apft = nullptr;
  void* stepmods = nullptr;
  void* pump = nullptr;
  void* cav = nullptr;
  void* ocm = nullptr;
  void* oc_spec = nullptr;
  void* esd_sum = nullptr;
  void* xesd_sum = nullptr;
  int esd_base = 0;  // trips_done when the run averages were last cleared
  void* hist = nullptr;
  void* sm_hist = nullptr;
  int hist_cap = 0;
  int hist_base = 0;  // trips_done when the history started
  void* trip_log = nullptr;
};

struct ofx_opo {
  ofx_context* ctx = nullptr;
  int prec = OFX_F32;
  size_t csz = 8;
  int n_inst = 0;
  int trip_cap = 0;
  std::vector<ofx_pass_params> pass;
  std::vector<InstDesc> inst_h;
  std::vector<InstBufs> bufs;
  std::vector<ClusterDesc> cl_h;
  std::vector<int> order_h, wgc_h, wgr_h;
  std::vector<int> inst_C, inst_n1;
  InstDesc* inst_d = nullptr;
  ClusterDesc* cl_d = nullptr;
  int* order_d = nullptr;
  int* wgc_d = nullptr;
  int* wgr_d = nullptr;
  int* xcc_d = nullptr;
  int* status_d = nullptr;
  int* trips_d = nullptr;
  int* fresh_d = nullptr;
  long long* counters_d = nullptr;
  void* twiddle = nullptr;
  int nmax = 0;
  int grid = 0, block = 256;
  size_t lds = 0;
  std::vector<void*> allocs;
  std::vector<unsigned*> zero_ptrs;
  std::vector<size_t> zero_words;
  unsigned* epoch_d = nullptr;  // [clusters][2] tag epochs (advanced before every launch)
  double wait_s = 10.0;         // exchange wait limit
  int xmode = 0;                // exchange mode of the kernel: 0 tagged polling, 1 counter barrier
  int graph_trips = -1;
  double graph_tol = 0.0;
};

namespace {

int fail(ofx_context* ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(ofx_context* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

int ilog2i(long long n) {
  int l = 0;
  while ((1LL << l) < n) l++;
  return ((1LL << l) == n) ? l : -1;
}

// Heun keeps A0, E and K in LDS with 2 halo columns; RK23 keeps A0, E, K1,
// K2, K3 with 4 (three stages deep, see rk23_pass).
int halo_of(int method) { return method == OFX_METHOD_RK23 ? 4 : 2; }
int nbuf_of(int method) { return method == OFX_METHOD_RK23 ? 5 : 3; }

size_t lds_bytes_for(int n1, int n2, int C, int method, size_t csz) {
  const int W = n1 / C, Rr = n2 / C, LC = W + 2 * halo_of(method);
  const size_t b1 = (size_t)LC * (size_t)(n2 + 1), b2 = (size_t)Rr * (size_t)(n1 + 1);
  // + the tables (2 n1 + 2 n2 complex), 16 bytes of alignment, the FP64
  // four-step factors (n1 + n2 complex doubles) and the own columns' absorber
  // factors (W n2 reals, half a complex each)
  return ((size_t)nbuf_of(method) * std::max(b1, b2) + (size_t)(2 * n1 + 2 * n2) +
          ((size_t)W * (size_t)n2 + 1) / 2) * csz + 16 + (size_t)(n1 + n2) * 16;
}

// largest cluster size the method allows (RK23 exchanges 3 boundary columns
// of K4 with each neighbour, so each workgroup owns at least 4 columns)
int max_cluster(int method, int n1, int n2) {
  return method == OFX_METHOD_RK23 ? std::min(n1 / 4, n2) : std::min(n1, n2);
}

int dev_alloc(ofx_opo* r, void** p, size_t bytes) {
  int st = be_malloc(r->ctx->be, p, bytes ? bytes : 16);
  if (st) return fail(r->ctx, OFX_E_NOMEM, "device allocation of %zu bytes failed: %s", bytes,
                      be_error(r->ctx->be));
  r->allocs.push_back(*p);
  return OFX_OK;
}

int upload(ofx_opo* r, void* dst, const void* src, size_t bytes) {
  if (be_h2d(r->ctx->be, dst, src, bytes))
    return fail(r->ctx, OFX_E_HIP, "host-to-device copy failed: %s", be_error(r->ctx->be));
  return OFX_OK;
}

// FP64 host array -> precision of the run (staging kept alive in `keep`)
const void* convert(const double* src, size_t count, int prec, std::vector<std::vector<char>>& keep) {
  keep.emplace_back(count * (prec == OFX_F64 ? 8 : 4));
  char* out = keep.back().data();
  if (prec == OFX_F64) {
    memcpy(out, src, count * 8);
  } else {
    float* f = (float*)out;
    for (size_t i = 0; i < count; i++) f[i] = (float)src[i];
  }
  return out;
}

// t[2m], t[2m+1] = exp(-2 pi i m / nmax), m < nmax; angles are dyadic
// fractions of a turn
void fill_twiddle(int nmax, double* t) {
  for (int m = 0; m < nmax; m++) {
    const double a = -2.0 * M_PI * (double)m / (double)nmax;
    t[2 * m] = cos(a);
    t[2 * m + 1] = sin(a);
  }
}

void* twiddle_table(ofx_context* ctx, int nmax, int prec) {
  auto key = std::make_pair(nmax, prec);
  auto it = ctx->tw.find(key);
  if (it != ctx->tw.end()) return it->second;
  std::vector<double> t(2 * (size_t)nmax);
  fill_twiddle(nmax, t.data());
  void* p = nullptr;
  const size_t bytes = (size_t)nmax * (prec == OFX_F64 ? 16 : 8);
  if (be_malloc(ctx->be, &p, bytes)) return nullptr;
  if (prec == OFX_F64) {
    be_h2d(ctx->be, p, t.data(), bytes);
  } else {
    std::vector<float> f(2 * (size_t)nmax);
    for (size_t i = 0; i < f.size(); i++) f[i] = (float)t[i];
    be_h2d(ctx->be, p, f.data(), bytes);
    be_sync(ctx->be);
  }
  be_sync(ctx->be);
  ctx->tw[key] = p;
  return p;
}

// x - 2*pi*round(x/(2*pi)), two-term split of 2*pi with fma (exact q*HI)
double reduce_2pi_host(double x) {
  const double TWO_PI_HI = 6.283185307179586232;
  const double TWO_PI_LO = 2.4492935982947064e-16;
  double q = rint(x * 0.15915494309189533577);
  return fma(-q, TWO_PI_LO, fma(-q, TWO_PI_HI, x));
}

// exp(i w0 t_j) and the absorbing-boundary factor for the fftshifted time
// axis t_j = fftshift((-N/2:N/2-1)*dt) (NEE_CPU.m:12-14, 53, 60, 114-121)
void time_tables(int n, double dt, double w0, double* wt, double* ab) {
  const double tr = (double)(n / 2 - 1) * dt - (double)(-(n / 2)) * dt;  // range(t)
  for (int j = 0; j < n; j++) {
    const double t = (double)(j < n / 2 ? j : j - n) * dt;
    const double th = reduce_2pi_host(w0 * t);
    wt[2 * (size_t)j] = cos(th);
    wt[2 * (size_t)j + 1] = sin(th);
    const double at2 = fabs(2.0 * t);
    double f = 1.0;
    if (tr - at2 < 0.25 * tr) {
      const double sh = 1.0 / cosh(24.0 * (1.0 - at2 / tr));
      f = 1.0 - sh * sh;
    }
    ab[(size_t)j] = f;
  }
}

int check_pass(ofx_context* ctx, const ofx_pass_params* p) {
  if (!p) return fail(ctx, OFX_E_ARG, "null pass parameters");
  const int lg = ilog2i(p->n);
  if (lg < 6 || lg > 16) return fail(ctx, OFX_E_ARG, "N=%d must be a power of two in [64, 65536]", p->n);
  if (p->nsteps < 1) return fail(ctx, OFX_E_ARG, "NSteps=%d must be >= 1", p->nsteps);
  if (!(p->sel > 1.0)) return fail(ctx, OFX_E_ARG, "sel=%g must be > 1", p->sel);
  if (!(p->dt > 0) || !(p->w0 > 0) || !(p->h > 0))
    return fail(ctx, OFX_E_ARG, "dt, w0 and h must be positive");
  if (!isfinite(p->bdiffw0) || !isfinite(p->max_err) || !isfinite(p->min_err))
    return fail(ctx, OFX_E_ARG, "bdiffw0, maxErr, minErr must be finite");
  if (p->n_cols < 0) return fail(ctx, OFX_E_ARG, "n_cols must be >= 0");
  if (p->method != OFX_METHOD_HEUN && p->method != OFX_METHOD_RK23)
    return fail(ctx, OFX_E_UNSUPPORTED, "method %d is not available on the device path", p->method);
  if (p->method == OFX_METHOD_RK23 && (int)p->sel > p->nsteps)
    return fail(ctx, OFX_E_ARG, "RK23: sel=%g exceeds NSteps=%d (no chunk)", p->sel, p->nsteps);
  return OFX_OK;
}

// Launch geometry of one instance of a batch of n_inst: the four-step split
// n1 (N = n1 n2) and the cluster size C (workgroups per instance).
int plan_instance(ofx_context* ctx, const ofx_pass_params& p, int n_inst, const ofx_tuning* tune,
                  int spec_cluster, int spec_n1, size_t csz, int* C_out, int* n1_out) {
  const int n = p.n, lg = ilog2i(n);
  int n1 = 1 << ((lg + 1) / 2);
  if (tune && tune->n1) n1 = tune->n1;
  if (spec_n1) n1 = spec_n1;
  if (ilog2i(n1) < 2 || n % n1 || n / n1 < 4)
    return fail(ctx, OFX_E_ARG, "n1=%d does not split N=%d into two powers of two >= 4", n1, n);
  const int n2 = n / n1;
  const int meth = p.method;
  const int cmax = max_cluster(meth, n1, n2);
  int cfit = 0;
  for (int C = 1; C <= cmax; C <<= 1) {
    if ((int)lds_bytes_for(n1, n2, C, meth, csz) <= ctx->info.lds_max - 1024) {
      cfit = C;
      break;
    }
  }
  if (!cfit)
    return fail(ctx, OFX_E_ARG, "N=%d does not fit the %d-byte LDS limit at any cluster size", n,
                ctx->info.lds_max);
  int C = cfit;
  const int want_C = spec_cluster ? spec_cluster : (tune ? tune->cluster : 0);
  if (want_C) {
    C = want_C;
    if (ilog2i(C) < 0 || C > cmax ||
        (int)lds_bytes_for(n1, n2, C, meth, csz) > ctx->info.lds_max - 1024)
      return fail(ctx, OFX_E_ARG, "cluster size %d is invalid for N=%d (n1=%d)", C, n, n1);
  } else {
    // Spread a small batch over more workgroups: a step's latency is set by
    // the per-workgroup share until it falls to ~256 points (measured on
    // N = 2^15, FP32 / FP64 us per adaptive step: 20.1 / 24.2 at 32
    // workgroups, 16.6 / 21.1 at 64, 15.5 / 19.6 at 128 with 256-thread
    // workgroups).
    const int slots_est = ctx->info.cus;
    while (C * 2 <= cmax && (long long)n_inst * C * 2 <= slots_est && C * 2 <= 128 &&
           n / (C * 2) >= 256)
      C *= 2;
  }
  *C_out = C;
  *n1_out = n1;
  return OFX_OK;
}

// Threads per workgroup: 256 once every cluster has 128 members or more (a
// share of <= 256 points per workgroup), 512 below.
int plan_block(ofx_context* ctx, const ofx_tuning* tune, int cmin, int* block) {
  const int exchange = tune ? tune->exchange : 0;
  if (exchange < 0 || exchange > 2)
    return fail(ctx, OFX_E_ARG, "exchange=%d must be 0 (auto), 1 (tags) or 2 (barrier)", exchange);
  *block = (tune && tune->block) ? tune->block : (cmin >= 128 ? 256 : 512);
  if (*block < 64 || *block > 512 || (*block & 63))
    return fail(ctx, OFX_E_ARG, "block=%d must be a multiple of 64 in [64,512]", *block);
  return OFX_OK;
}

// Exchange mode of a launch: polling the tagged words wins when every
// cluster has at least 64 workgroups (a few elements per thread, latency
// bound); with larger shares per workgroup the plain words plus a counter
// barrier move half the bytes and win (measured on N = 2^15).
int exchange_mode(int exchange, int cmin) {
  return exchange == 1 ? 0 : exchange == 2 ? 1 : (cmin >= 64 ? 0 : 1);
}

}  // namespace

extern "C" {

int ofx_create(int device, ofx_context** out) {
  if (!out) return OFX_E_ARG;
  *out = nullptr;
  ofx_context* ctx = new ofx_context();
  int st = be_create(device, &ctx->be, &ctx->info);
  if (st) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return OFX_OK;
}

void ofx_destroy(ofx_context* ctx) {
  if (!ctx) return;
  ofx_opo_destroy(ctx->pass_run);
  ctx->pass_run = nullptr;
  for (auto& kv : ctx->tw) be_free(ctx->be, kv.second);
  be_destroy(ctx->be);
  delete ctx;
}

const char* ofx_last_error(const ofx_context* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int ofx_device_info(const ofx_context* ctx, char* name, int name_len, int* cus, int* wallclock_khz) {
  if (!ctx) return OFX_E_ARG;
  if (name && name_len > 0) {
    strncpy(name, ctx->info.name, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cus) *cus = ctx->info.cus;
  if (wallclock_khz) *wallclock_khz = ctx->info.wallclock_khz;
  return OFX_OK;
}

void ofx_opo_destroy(ofx_opo* r) {
  if (!r) return;
  be_sync(r->ctx->be);
  be_forget_graph(r->ctx->be, r);
  for (void* p : r->allocs) be_free(r->ctx->be, p);
  delete r;
}

int ofx_opo_create(ofx_context* ctx, int precision, int n_inst, const ofx_opo_spec* specs,
                   int trip_cap, const ofx_tuning* tune, ofx_opo** out) {
  if (!ctx || !out) return OFX_E_ARG;
  *out = nullptr;
  if (precision != OFX_F32 && precision != OFX_F64)
    return fail(ctx, OFX_E_ARG, "precision must be OFX_F32 or OFX_F64");
  if (n_inst < 1 || !specs) return fail(ctx, OFX_E_ARG, "need at least one instance");
  if (trip_cap < 0) return fail(ctx, OFX_E_ARG, "trip_cap must be >= 0");
  for (int i = 0; i < n_inst; i++) {
    int st = check_pass(ctx, &specs[i].pass);
    if (st) return st;
    const ofx_opo_spec& s = specs[i];
    if (!s.beta || !s.alpha || !s.G || !s.field)
      return fail(ctx, OFX_E_ARG, "instance %d: beta, alpha, G and field are required", i);
    if (s.cav && !s.ocm) return fail(ctx, OFX_E_ARG, "instance %d: cav given without ocm", i);
    if (!isfinite(s.delay)) return fail(ctx, OFX_E_ARG, "instance %d: delay must be finite", i);
  }
  ofx_opo* r = new ofx_opo();
  r->ctx = ctx;
  r->prec = precision;
  r->csz = precision == OFX_F64 ? 16 : 8;
  r->n_inst = n_inst;
  r->trip_cap = trip_cap;
  const size_t rsz = precision == OFX_F64 ? 8 : 4;
  std::vector<std::vector<char>> keep;
  int st = OFX_OK;

  // ---- launch geometry per instance ----
  r->inst_C.resize(n_inst);
  r->inst_n1.resize(n_inst);
  int nmax = 0;
  for (int i = 0; i < n_inst; i++) {
    nmax = std::max(nmax, specs[i].pass.n);
    st = plan_instance(ctx, specs[i].pass, n_inst, tune, specs[i].cluster, specs[i].n1, r->csz,
                       &r->inst_C[i], &r->inst_n1[i]);
    if (st) {
      delete r;
      return st;
    }
  }
  r->nmax = nmax;
  int cmin0 = 1 << 30;
  for (int i = 0; i < n_inst; i++) cmin0 = std::min(cmin0, r->inst_C[i]);
  const int exchange = tune ? tune->exchange : 0;
  st = plan_block(ctx, tune, cmin0, &r->block);
  if (st) {
    delete r;
    return st;
  }
  size_t lds = 0;
  int cmin = 1 << 30;
  for (int i = 0; i < n_inst; i++) {
    const int n1 = r->inst_n1[i], n2 = specs[i].pass.n / n1;
    lds = std::max(lds, lds_bytes_for(n1, n2, r->inst_C[i], specs[i].pass.method, r->csz));
    cmin = std::min(cmin, r->inst_C[i]);
  }
  r->lds = lds;
  r->xmode = exchange_mode(exchange, cmin);
  bool all_rk23 = n_inst > 0;
  for (int i = 0; i < n_inst; i++) all_rk23 = all_rk23 && specs[i].pass.method == OFX_METHOD_RK23;
  if (all_rk23) r->xmode |= XMODE_RK23_LDS;
  const int bpc = be_blocks_per_cu(ctx->be, precision, r->xmode, r->block, lds);
  if (bpc < 1) {
    delete r;
    return fail(ctx, OFX_E_ARG, "kernel cannot launch with block=%d and %zu bytes of LDS", r->block, lds);
  }
  const int slots = ctx->info.cus * bpc;

  // ---- clusters: group instances by (C, n, n1, method); LPT over cost ----
  std::map<std::tuple<int, int, int, int>, std::vector<int>> groups;
  for (int i = 0; i < n_inst; i++)
    groups[std::make_tuple(r->inst_C[i], specs[i].pass.n, r->inst_n1[i], specs[i].pass.method)]
        .push_back(i);
  long long want = 0;
  for (auto& g : groups) want += (long long)std::get<0>(g.first) * (long long)g.second.size();
  std::vector<std::vector<int>> cl_members;
  std::vector<int> cl_C, cl_method;
  for (auto& g : groups) {
    const int C = std::get<0>(g.first);
    std::vector<int> ids = g.second;
    long long ncl = (long long)ids.size();
    if (want > slots) ncl = std::max(1LL, (long long)ids.size() * slots / want);
    ncl = std::min(ncl, (long long)std::max(1, slots / C));
    std::sort(ids.begin(), ids.end(), [&](int a, int b) {
      const double ca = (double)specs[a].pass.nsteps, cb = (double)specs[b].pass.nsteps;
      return ca != cb ? ca > cb : a < b;
    });
    std::vector<std::vector<int>> mem((size_t)ncl);
    std::vector<double> load((size_t)ncl, 0.0);
    for (int id : ids) {
      size_t best = 0;
      for (size_t c = 1; c < load.size(); c++)
        if (load[c] < load[best]) best = c;
      mem[best].push_back(id);
      load[best] += (double)specs[id].pass.nsteps;
    }
    for (auto& m : mem) {
      if (m.empty()) continue;
      cl_members.push_back(m);
      cl_C.push_back(C);
      cl_method.push_back(std::get<3>(g.first));
    }
  }
  long long used = 0;
  for (int C : cl_C) used += C;
  if (used > slots) {
    delete r;
    return fail(ctx, OFX_E_ARG, "schedule needs %lld workgroups, only %d fit at once", used, slots);
  }

  // ---- XCD-aware placement: workgroup id g runs on XCD g % nxcd ----
  const int nx = std::max(1, ctx->info.nxcd);
  const int per_x = slots / nx;
  std::vector<int> freex(nx, per_x);
  std::vector<int> order_cl(cl_C.size());
  for (size_t i = 0; i < order_cl.size(); i++) order_cl[i] = (int)i;
  std::stable_sort(order_cl.begin(), order_cl.end(), [&](int a, int b) { return cl_C[a] > cl_C[b]; });
  std::vector<std::vector<int>> cl_wgs(cl_C.size());
  int maxid = -1;
  for (int ci : order_cl) {
    const int C = cl_C[ci];
    int x = (int)(std::max_element(freex.begin(), freex.end()) - freex.begin());
    if (freex[x] >= C) {
      for (int q = 0; q < C; q++) {
        const int m = per_x - freex[x];
        cl_wgs[ci].push_back(x + nx * m);
        freex[x]--;
      }
    } else {
      for (int q = 0; q < C; q++) {
        x = (int)(std::max_element(freex.begin(), freex.end()) - freex.begin());
        const int m = per_x - freex[x];
        cl_wgs[ci].push_back(x + nx * m);
        freex[x]--;
      }
    }
    for (int id : cl_wgs[ci]) maxid = std::max(maxid, id);
  }
  r->grid = maxid + 1;
  r->wgc_h.assign(r->grid, -1);
  r->wgr_h.assign(r->grid, 0);

  // ---- device buffers ----
  st = dev_alloc(r, (void**)&r->status_d, sizeof(int) * n_inst);
  if (!st) st = dev_alloc(r, (void**)&r->trips_d, sizeof(int) * n_inst);
  if (!st) st = dev_alloc(r, (void**)&r->fresh_d, sizeof(int) * n_inst);
  if (!st) st = dev_alloc(r, (void**)&r->counters_d, sizeof(long long) * 2 * n_inst);
  if (!st) st = dev_alloc(r, (void**)&r->xcc_d, sizeof(int) * r->grid);
  if (st) { ofx_opo_destroy(r); return st; }
  {
    std::vector<char> z(std::max(sizeof(long long) * 2 * n_inst, sizeof(int) * (size_t)r->grid), 0);
    upload(r, r->status_d, z.data(), sizeof(int) * n_inst);
    upload(r, r->trips_d, z.data(), sizeof(int) * n_inst);
    upload(r, r->fresh_d, z.data(), sizeof(int) * n_inst);
    upload(r, r->counters_d, z.data(), sizeof(long long) * 2 * n_inst);
    upload(r, r->xcc_d, z.data(), sizeof(int) * r->grid);
    be_sync(ctx->be);
  }
  r->twiddle = twiddle_table(ctx, nmax, precision);
  if (!r->twiddle) { ofx_opo_destroy(r); return fail(ctx, OFX_E_NOMEM, "twiddle table allocation failed"); }

  r->pass.resize(n_inst);
  r->inst_h.resize(n_inst);
  r->bufs.resize(n_inst);
  for (int i = 0; i < n_inst && !st; i++) {
    const ofx_opo_spec& s = specs[i];
    const ofx_pass_params& p = s.pass;
    r->pass[i] = p;
    InstBufs& b = r->bufs[i];
    const size_t n = (size_t)p.n;
    st = dev_alloc(r, &b.field, n * r->csz);
    if (!st) st = dev_alloc(r, &b.wt, n * r->csz);
    if (!st) st = dev_alloc(r, &b.absorb, n * rsz);
    if (!st) st = dev_alloc(r, &b.beta, n * 8);
    if (!st) st = dev_alloc(r, &b.alpha, n * 8);
    if (!st) st = dev_alloc(r, &b.G, (size_t)p.nsteps * 8);
    if (!st) st = dev_alloc(r, &b.stepmods, (size_t)p.nsteps * 4);
    if (!st && p.n_cols > 0) st = dev_alloc(r, &b.apft, n * (size_t)p.n_cols * r->csz);
    if (!st && s.cav) {
      st = dev_alloc(r, &b.cav, n * r->csz);
      if (!st) st = dev_alloc(r, &b.ocm, n * r->csz);
      if (!st) st = dev_alloc(r, &b.oc_spec, n * r->csz);
      if (!st) st = dev_alloc(r, &b.esd_sum, n * 8);
      if (!st) st = dev_alloc(r, &b.xesd_sum, n * 8);
      if (!st && s.pump) st = dev_alloc(r, &b.pump, n * r->csz);
    }
    if (!st && trip_cap > 0) st = dev_alloc(r, &b.trip_log, (size_t)trip_cap * 4 * 8);
    if (st) break;
    upload(r, b.field, convert(s.field, 2 * n, precision, keep), n * r->csz);
    {
      std::vector<double> wt(2 * n), ab(n);
      time_tables(p.n, p.dt, p.w0, wt.data(), ab.data());
      upload(r, b.wt, convert(wt.data(), wt.size(), precision, keep), n * r->csz);
      upload(r, b.absorb, convert(ab.data(), ab.size(), precision, keep), n * rsz);
    }
    upload(r, b.beta, s.beta, n * 8);
    upload(r, b.alpha, s.alpha, n * 8);
    upload(r, b.G, s.G, (size_t)p.nsteps * 8);
    {
      keep.emplace_back((size_t)p.nsteps * 4, 0);
      upload(r, b.stepmods, keep.back().data(), (size_t)p.nsteps * 4);
    }
    if (b.apft) {
      keep.emplace_back(n * (size_t)p.n_cols * r->csz, 0);
      upload(r, b.apft, keep.back().data(), n * (size_t)p.n_cols * r->csz);
    }
    if (b.cav) {
      upload(r, b.cav, convert(s.cav, 2 * n, precision, keep), n * r->csz);
      upload(r, b.ocm, convert(s.ocm, 2 * n, precision, keep), n * r->csz);
      keep.emplace_back(n * r->csz, 0);
      upload(r, b.oc_spec, keep.back().data(), n * r->csz);
      keep.emplace_back(n * 8, 0);
      upload(r, b.esd_sum, keep.back().data(), n * 8);
      upload(r, b.xesd_sum, keep.back().data(), n * 8);
      if (b.pump) upload(r, b.pump, convert(s.pump, 2 * n, precision, keep), n * r->csz);
    }
    if (b.trip_log) {
      keep.emplace_back((size_t)trip_cap * 4 * 8, 0);
      upload(r, b.trip_log, keep.back().data(), (size_t)trip_cap * 4 * 8);
    }
    InstDesc& d = r->inst_h[i];
    memset(&d, 0, sizeof d);
    d.n = p.n;
    d.n1 = r->inst_n1[i];
    d.n2 = p.n / d.n1;
    d.lg1 = ilog2i(d.n1);
    d.lg2 = ilog2i(d.n2);
    d.nsteps = p.nsteps;
    d.n_cols = p.n_cols;
    d.method = p.method;
    d.w0 = p.w0;
    d.bdiffw0 = p.bdiffw0;
    d.h = p.h;
    d.dt = p.dt;
    d.max_err = p.max_err;
    d.min_err = p.min_err;
    d.sel = p.sel;
    d.delay = s.delay;
    d.beta = (const double*)b.beta;
    d.alpha = (const double*)b.alpha;
    d.G = (const double*)b.G;
    d.field = b.field;
    d.wt = b.wt;
    d.absorb = b.absorb;
    d.apft = b.apft;
    d.stepmods = (float*)b.stepmods;
    d.pump = b.pump;
    d.cav = b.cav;
    d.ocm = b.ocm;
    d.oc_spec = b.oc_spec;
    d.esd_sum = (double*)b.esd_sum;
    d.xesd_sum = (double*)b.xesd_sum;
    d.trip_log = (double*)b.trip_log;
    d.trip_cap = trip_cap;
    d.status = r->status_d + i;
    d.trips_done = r->trips_d + i;
    d.fresh = r->fresh_d + i;
    d.counters = r->counters_d + 2 * i;
  }
  if (st) { be_sync(ctx->be); ofx_opo_destroy(r); return st; }

  // clusters; exchange buffers hold tagged elements (twice the complex size)
  // in two parity halves, zeroed at allocation (tag 0 is never awaited)
  r->cl_h.resize(cl_C.size());
  st = dev_alloc(r, (void**)&r->epoch_d, sizeof(unsigned) * 2 * cl_C.size());
  for (size_t c = 0; c < cl_C.size() && !st; c++) {
    ClusterDesc& cd = r->cl_h[c];
    memset(&cd, 0, sizeof cd);
    cd.C = cl_C[c];
    cd.first_inst = (int)r->order_h.size();
    cd.n_inst = (int)cl_members[c].size();
    cd.halo = halo_of(cl_method[c]);
    int nc = 0, n2c = 0;
    for (int id : cl_members[c]) {
      r->order_h.push_back(id);
      nc = std::max(nc, specs[id].pass.n);
      n2c = std::max(n2c, specs[id].pass.n / r->inst_n1[id]);
    }
    const size_t tsz = 2 * r->csz;  // bytes of one tagged complex element
    st = dev_alloc(r, &cd.t1, 2 * (size_t)nc * tsz);
    if (!st) st = dev_alloc(r, &cd.t2, 2 * (size_t)nc * tsz);
    if (!st) st = dev_alloc(r, &cd.slots, 2 * 8 * 4 * (size_t)cd.C * tsz);
    if (!st) st = dev_alloc(r, (void**)&cd.bar, 64);
    if (!st) st = dev_alloc(r, (void**)&cd.abort_flag, 64);
    if (!st) st = dev_alloc(r, (void**)&cd.xcc, sizeof(int) * (size_t)cd.C);
    if (!st && cl_method[c] == OFX_METHOD_RK23)
      st = dev_alloc(r, &cd.kx, 2 * (size_t)cd.C * 6 * (size_t)n2c * tsz);
    if (st) break;
    cd.epoch = r->epoch_d + 2 * c;
    r->zero_ptrs.push_back(cd.bar);
    r->zero_words.push_back(1);
    r->zero_ptrs.push_back((unsigned*)cd.abort_flag);
    r->zero_words.push_back(1);
    for (int q = 0; q < cd.C; q++) {
      r->wgc_h[cl_wgs[c][q]] = (int)c;
      r->wgr_h[cl_wgs[c][q]] = q;
    }
  }
  if (!st) st = dev_alloc(r, (void**)&r->inst_d, sizeof(InstDesc) * n_inst);
  if (!st) st = dev_alloc(r, (void**)&r->cl_d, sizeof(ClusterDesc) * r->cl_h.size());
  if (!st) st = dev_alloc(r, (void**)&r->order_d, sizeof(int) * r->order_h.size());
  if (!st) st = dev_alloc(r, (void**)&r->wgc_d, sizeof(int) * r->grid);
  if (!st) st = dev_alloc(r, (void**)&r->wgr_d, sizeof(int) * r->grid);
  if (st) { be_sync(ctx->be); ofx_opo_destroy(r); return st; }
  upload(r, r->inst_d, r->inst_h.data(), sizeof(InstDesc) * n_inst);
  upload(r, r->cl_d, r->cl_h.data(), sizeof(ClusterDesc) * r->cl_h.size());
  upload(r, r->order_d, r->order_h.data(), sizeof(int) * r->order_h.size());
  upload(r, r->wgc_d, r->wgc_h.data(), sizeof(int) * r->grid);
  upload(r, r->wgr_d, r->wgr_h.data(), sizeof(int) * r->grid);
  if (be_sync(ctx->be)) {
    st = fail(ctx, OFX_E_HIP, "upload failed: %s", be_error(ctx->be));
    ofx_opo_destroy(r);
    return st;
  }
  *out = r;
  return OFX_OK;
}

static LaunchArgs make_args(ofx_opo* r, int trips, double tol) {
  LaunchArgs a;
  memset(&a, 0, sizeof a);
  a.inst = r->inst_d;
  a.clusters = r->cl_d;
  a.inst_order = r->order_d;
  a.wg_cluster = r->wgc_d;
  a.wg_rank = r->wgr_d;
  a.twiddle = r->twiddle;
  a.nmax = r->nmax;
  a.trips = trips;
  a.conv_tol = tol;
  // exchange wait limit in wall-clock ticks
  a.timeout_ticks = (unsigned long long)((double)r->ctx->info.wallclock_khz * 1000.0 * r->wait_s);
  a.xcc_ids = r->xcc_d;
  return a;
}

int ofx_opo_launch(ofx_opo* r, int trips, double conv_tol) {
  if (!r) return OFX_E_ARG;
  if (trips < 0) return fail(r->ctx, OFX_E_ARG, "trips must be >= 0");
  LaunchArgs a = make_args(r, trips, conv_tol);
  int st = be_launch(r->ctx->be, a, r->prec, r->xmode, r->grid, r->block, r->lds,
                     r->zero_ptrs.data(), r->zero_words.data(), (int)r->zero_ptrs.size(),
                     r->epoch_d, (int)r->cl_h.size(), nullptr, 0);
  if (st) return fail(r->ctx, OFX_E_HIP, "launch failed: %s", be_error(r->ctx->be));
  return OFX_OK;
}

int ofx_opo_graph_replay(ofx_opo* r, int trips, double conv_tol, int replays) {
  if (!r || replays < 1 || trips < 0) return OFX_E_ARG;
  if (r->graph_trips != trips || r->graph_tol != conv_tol) {
    be_forget_graph(r->ctx->be, r);
    r->graph_trips = trips;
    r->graph_tol = conv_tol;
  }
  LaunchArgs a = make_args(r, trips, conv_tol);
  int st = be_launch(r->ctx->be, a, r->prec, r->xmode, r->grid, r->block, r->lds,
                     r->zero_ptrs.data(), r->zero_words.data(), (int)r->zero_ptrs.size(),
                     r->epoch_d, (int)r->cl_h.size(), r, replays);
  if (st) return fail(r->ctx, OFX_E_HIP, "graph replay failed: %s", be_error(r->ctx->be));
  return OFX_OK;
}

int ofx_opo_set_wait_limit(ofx_opo* r, double seconds) {
  if (!r) return OFX_E_ARG;
  if (!(seconds >= 0.0 && seconds <= 1e6))
    return fail(r->ctx, OFX_E_ARG, "wait limit must be in [0, 1e6] s");
  if (seconds != r->wait_s) {
    be_forget_graph(r->ctx->be, r);  // the captured launch holds the old limit
    r->graph_trips = -1;
    r->wait_s = seconds;
  }
  return OFX_OK;
}

int ofx_opo_sync(ofx_opo* r) {
  if (!r) return OFX_E_ARG;
  if (be_sync(r->ctx->be)) return fail(r->ctx, OFX_E_HIP, "synchronize failed: %s", be_error(r->ctx->be));
  std::vector<int> stv(r->n_inst);
  if (be_d2h(r->ctx->be, stv.data(), r->status_d, sizeof(int) * r->n_inst) || be_sync(r->ctx->be))
    return fail(r->ctx, OFX_E_HIP, "status readback failed: %s", be_error(r->ctx->be));
  for (int i = 0; i < r->n_inst; i++) {
    if (stv[i] == ST_BARRIER_TIMEOUT)
      return fail(r->ctx, OFX_E_TIMEOUT, "instance %d: a cluster member timed out waiting for exchange data", i);
  }
  for (int i = 0; i < r->n_inst; i++) {
    if (stv[i] == ST_NONFINITE)
      return fail(r->ctx, OFX_E_NONFINITE, "instance %d: field became non-finite", i);
  }
  return OFX_OK;
}

int ofx_opo_last_ms(ofx_opo* r, double* ms) {
  if (!r || !ms) return OFX_E_ARG;
  if (be_last_ms(r->ctx->be, ms)) return fail(r->ctx, OFX_E_HIP, "no timing available");
  return OFX_OK;
}

static int read_conv(ofx_opo* r, void* src, double* dst, size_t count) {
  std::vector<char> tmp(count * (r->prec == OFX_F64 ? 8 : 4));
  if (be_d2h(r->ctx->be, tmp.data(), src, tmp.size()) || be_sync(r->ctx->be))
    return fail(r->ctx, OFX_E_HIP, "readback failed: %s", be_error(r->ctx->be));
  if (r->prec == OFX_F64) {
    memcpy(dst, tmp.data(), count * 8);
  } else {
    const float* f = (const float*)tmp.data();
    for (size_t i = 0; i < count; i++) dst[i] = f[i];
  }
  return OFX_OK;
}

int ofx_opo_read(ofx_opo* r, int inst, double* field, double* oc_spec, double* trip_log, int* status,
                 int* trips_done, long long* counters, double* apft, double* stepmods) {
  if (!r || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  if (be_sync(r->ctx->be)) return fail(r->ctx, OFX_E_HIP, "synchronize failed: %s", be_error(r->ctx->be));
  const InstBufs& b = r->bufs[inst];
  const ofx_pass_params& p = r->pass[inst];
  const size_t n = (size_t)p.n;
  int st = OFX_OK;
  if (field) st = read_conv(r, b.field, field, 2 * n);
  if (!st && oc_spec) {
    if (!b.oc_spec) return fail(r->ctx, OFX_E_ARG, "instance %d has no cavity", inst);
    st = read_conv(r, b.oc_spec, oc_spec, 2 * n);
  }
  if (!st && trip_log && b.trip_log) {
    if (be_d2h(r->ctx->be, trip_log, b.trip_log, (size_t)r->trip_cap * 32) || be_sync(r->ctx->be))
      st = fail(r->ctx, OFX_E_HIP, "readback failed");
  }
  if (!st && status) {
    if (be_d2h(r->ctx->be, status, r->status_d + inst, sizeof(int)) || be_sync(r->ctx->be))
      st = fail(r->ctx, OFX_E_HIP, "readback failed");
  }
  if (!st && trips_done) {
    if (be_d2h(r->ctx->be, trips_done, r->trips_d + inst, sizeof(int)) || be_sync(r->ctx->be))
      st = fail(r->ctx, OFX_E_HIP, "readback failed");
  }
  if (!st && counters) {
    if (be_d2h(r->ctx->be, counters, r->counters_d + 2 * inst, 2 * sizeof(long long)) ||
        be_sync(r->ctx->be))
      st = fail(r->ctx, OFX_E_HIP, "readback failed");
  }
  if (!st && apft && b.apft) st = read_conv(r, b.apft, apft, 2 * n * (size_t)p.n_cols);
  if (!st && stepmods) {
    std::vector<float> sm((size_t)p.nsteps);
    if (be_d2h(r->ctx->be, sm.data(), b.stepmods, sm.size() * 4) || be_sync(r->ctx->be))
      st = fail(r->ctx, OFX_E_HIP, "readback failed");
    for (size_t j = 0; !st && j < sm.size(); j++) stepmods[j] = sm[j];
  }
  return st;
}

static int trips_done_of(ofx_opo* r, int inst, int* done) {
  if (be_sync(r->ctx->be) || be_d2h(r->ctx->be, done, r->trips_d + inst, sizeof(int)) ||
      be_sync(r->ctx->be))
    return fail(r->ctx, OFX_E_HIP, "readback failed: %s", be_error(r->ctx->be));
  return OFX_OK;
}

int ofx_opo_read_average(ofx_opo* r, int inst, double* oc_esd, double* xout_esd, int* trips) {
  if (!r || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  const InstBufs& b = r->bufs[inst];
  if (!b.esd_sum) return fail(r->ctx, OFX_E_ARG, "instance %d has no cavity", inst);
  int done = 0;
  int st = trips_done_of(r, inst, &done);
  if (st) return st;
  const int k = done - b.esd_base;
  if (trips) *trips = k;
  const size_t n = (size_t)r->pass[inst].n;
  void* src[2] = {b.esd_sum, b.xesd_sum};
  double* dst[2] = {oc_esd, xout_esd};
  for (int q = 0; q < 2; q++) {
    if (!dst[q]) continue;
    if (be_d2h(r->ctx->be, dst[q], src[q], n * 8) || be_sync(r->ctx->be))
      return fail(r->ctx, OFX_E_HIP, "readback failed: %s", be_error(r->ctx->be));
    for (size_t j = 0; j < n; j++) dst[q][j] = k > 0 ? dst[q][j] / k : 0.0;
  }
  return OFX_OK;
}

int ofx_opo_reset_average(ofx_opo* r, int inst) {
  if (!r || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  InstBufs& b = r->bufs[inst];
  if (!b.esd_sum) return fail(r->ctx, OFX_E_ARG, "instance %d has no cavity", inst);
  int done = 0;
  int st = trips_done_of(r, inst, &done);
  if (st) return st;
  const size_t n = (size_t)r->pass[inst].n;
  std::vector<double> z(n, 0.0);
  st = upload(r, b.esd_sum, z.data(), n * 8);
  if (!st) st = upload(r, b.xesd_sum, z.data(), n * 8);
  if (!st && be_sync(r->ctx->be)) st = fail(r->ctx, OFX_E_HIP, "upload failed");
  if (!st) b.esd_base = done;
  return st;
}

int ofx_opo_keep_history(ofx_opo* r, int trips) {
  if (!r || trips < 1) return OFX_E_ARG;
  for (int i = 0; i < r->n_inst; i++)
    if (r->bufs[i].hist) return fail(r->ctx, OFX_E_ARG, "the run already keeps a history");
  int st = OFX_OK;
  for (int i = 0; i < r->n_inst && !st; i++) {
    InstBufs& b = r->bufs[i];
    if (!b.esd_sum) continue;

    const size_t n = (size_t)r->pass[i].n;
    st = trips_done_of(r, i, &b.hist_base);
    if (!st) st = dev_alloc(r, &b.hist, (size_t)trips * n * r->csz);
    if (!st) st = dev_alloc(r, &b.sm_hist, (size_t)trips * (size_t)r->pass[i].nsteps * 4);
    if (!st) {
      b.hist_cap = trips;
      r->inst_h[i].hist = b.hist;
      r->inst_h[i].sm_hist = (float*)b.sm_hist;
      r->inst_h[i].hist_cap = trips;
    }
  }
  if (!st) st = upload(r, r->inst_d, r->inst_h.data(), sizeof(InstDesc) * r->n_inst);
  if (!st && be_sync(r->ctx->be)) st = fail(r->ctx, OFX_E_HIP, "upload failed");
  if (st) {
    // leave no instance half set up (the buffers stay with the run until it is destroyed)
    for (int i = 0; i < r->n_inst; i++) {
      r->bufs[i].hist = r->bufs[i].sm_hist = nullptr;
      r->bufs[i].hist_cap = 0;
      r->inst_h[i].hist = nullptr;
      r->inst_h[i].sm_hist = nullptr;
      r->inst_h[i].hist_cap = 0;
    }
  }
  return st;
}

int ofx_opo_read_history(ofx_opo* r, int inst, int trip, double* spectrum, double* stepmods) {
  if (!r || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  const InstBufs& b = r->bufs[inst];
  if (!b.hist) return fail(r->ctx, OFX_E_ARG, "instance %d keeps no history", inst);
  int done = 0;
  int st = trips_done_of(r, inst, &done);
  if (st) return st;
  const int lo = std::max(b.hist_base, done - b.hist_cap);
  if (trip < lo || trip >= done)
    return fail(r->ctx, OFX_E_ARG, "trip %d of instance %d is not held (held: %d..%d)", trip, inst,
                lo, done - 1);
  const size_t n = (size_t)r->pass[inst].n;
  if (spectrum) {
    st = read_conv(r, (char*)b.hist + (size_t)(trip % b.hist_cap) * n * r->csz, spectrum, 2 * n);
    if (st) return st;
  }
  if (stepmods) {
    const size_t ns = (size_t)r->pass[inst].nsteps;
    std::vector<float> sm(ns);
    if (be_d2h(r->ctx->be, sm.data(), (float*)b.sm_hist + (size_t)(trip % b.hist_cap) * ns,
               ns * 4) || be_sync(r->ctx->be))
      return fail(r->ctx, OFX_E_HIP, "readback failed: %s", be_error(r->ctx->be));
    for (size_t j = 0; j < ns; j++) stepmods[j] = sm[j];
  }
  return OFX_OK;
}

int ofx_opo_write_field(ofx_opo* r, int inst, const double* field) {
  if (!r || !field || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  std::vector<std::vector<char>> keep;
  const size_t n = (size_t)r->pass[inst].n;
  int st = upload(r, r->bufs[inst].field, convert(field, 2 * n, r->prec, keep), n * r->csz);
  const int running = ST_RUNNING, one = 1;
  if (!st) st = upload(r, r->status_d + inst, &running, sizeof(int));
  if (!st) st = upload(r, r->fresh_d + inst, &one, sizeof(int));
  if (!st && be_sync(r->ctx->be)) st = fail(r->ctx, OFX_E_HIP, "upload failed");
  return st;
}

int ofx_opo_geometry(const ofx_opo* r, int inst, int* cluster, int* n1, int* block, int* grid,
                     int* lds_bytes) {
  if (!r || inst < 0 || inst >= r->n_inst) return OFX_E_ARG;
  if (cluster) *cluster = r->inst_C[inst];
  if (n1) *n1 = r->inst_n1[inst];
  if (block) *block = r->block;
  if (grid) *grid = r->grid;
  if (lds_bytes) *lds_bytes = (int)r->lds;
  return OFX_OK;
}

int ofx_opo_xcc_ids(ofx_opo* r, int* ids, int cap, int* count) {
  if (!r || !ids || cap < 0) return OFX_E_ARG;
  const int n = std::min(cap, r->grid);
  if (be_d2h(r->ctx->be, ids, r->xcc_d, sizeof(int) * (size_t)n) || be_sync(r->ctx->be))
    return fail(r->ctx, OFX_E_HIP, "readback failed");
  if (count) *count = r->grid;
  return OFX_OK;
}

}  // extern "C"

// ---------------- host single pass (OPOmexBatch computation) ----------------
// OpticalSim.run calls the solver once per round trip with the same shapes
// and tables (OpticalSim.m:237-252), so the context keeps the device run of
// the last call: a call with the same precision, pass parameters, batch,
// G layout, outputs and tuning re-uploads the field, tables, ApFT and
// stepMods into it instead of allocating and planning a new one (the
// reference creates and destroys its buffers and FFT plan per call,
// OPO_CUDA_BATCH_XTAL.cu:415-419,518-522).
static PassKey pass_key_of(int prec, const ofx_pass_params* p, int batch, int g_batch, bool apft,
                           bool sm, const ofx_tuning* tune) {
  PassKey k;
  memset(&k, 0, sizeof k);  // padding too: keys are compared bytewise
  k.prec = prec;
  k.batch = batch;
  k.g_batch = g_batch;
  k.apft = apft ? 1 : 0;
  k.stepmods = sm ? 1 : 0;
  k.p.n = p->n;
  k.p.nsteps = p->nsteps;
  k.p.w0 = p->w0;
  k.p.bdiffw0 = p->bdiffw0;
  k.p.h = p->h;
  k.p.dt = p->dt;
  k.p.max_err = p->max_err;
  k.p.min_err = p->min_err;
  k.p.sel = p->sel;
  k.p.n_cols = apft ? p->n_cols : 0;
  k.p.method = p->method;
  if (tune) {
    k.has_tune = 1;
    k.t.cluster = tune->cluster;
    k.t.n1 = tune->n1;
    k.t.block = tune->block;
    k.t.exchange = tune->exchange;
  }
  return k;
}

template <typename T>
static int pass_host_impl(ofx_context* ctx, int prec, const ofx_pass_params* p, int batch, T* Et,
                          const T* alpha_w, const T* G, int g_batch, const T* beta_op, T* ApFT,
                          T* stepMods, int* status, const ofx_tuning* tune) {
  if (!ctx) return OFX_E_ARG;
  int st = check_pass(ctx, p);
  if (st) return st;
  if (batch == 0) return OFX_OK;  // an empty batch: nothing to solve
  if (batch < 0 || !Et || !alpha_w || !G || !beta_op)
    return fail(ctx, OFX_E_ARG, "Et, alpha_w, G, beta_op are required and batch >= 0");
  if (g_batch != 1 && g_batch != batch)
    return fail(ctx, OFX_E_ARG, "G must hold NSteps or NSteps*batch values (g_batch=%d)", g_batch);
  const size_t n = (size_t)p->n;
  const bool want_ap = ApFT && p->n_cols > 0;
  std::vector<double> beta(n), alpha(n), field(2 * n * batch), Gd((size_t)p->nsteps * g_batch);
  for (size_t j = 0; j < n; j++) {
    beta[j] = (double)beta_op[j];
    alpha[j] = (double)alpha_w[j];
  }
  for (size_t j = 0; j < field.size(); j++) field[j] = (double)Et[j];
  for (size_t j = 0; j < Gd.size(); j++) Gd[j] = (double)G[j];
  const PassKey key = pass_key_of(prec, p, batch, g_batch, want_ap, stepMods != nullptr, tune);
  ofx_opo* r = ctx->pass_run;
  std::vector<std::vector<char>> keep;
  if (r && memcmp(&key, &ctx->pass_key, sizeof key) == 0) {
    // same shapes: new data into the kept buffers (the time tables depend
    // only on N, dt and w0, which are part of the key)
    for (int b = 0; b < batch && !st; b++) {
      InstBufs& bf = r->bufs[b];
      st = upload(r, bf.field, convert(field.data() + 2 * n * b, 2 * n, prec, keep), n * r->csz);
      if (!st) st = upload(r, bf.beta, beta.data(), n * 8);
      if (!st) st = upload(r, bf.alpha, alpha.data(), n * 8);
      if (!st)
        st = upload(r, bf.G, Gd.data() + (g_batch == 1 ? 0 : (size_t)b * p->nsteps),
                    (size_t)p->nsteps * 8);
    }
  } else {
    ofx_opo_destroy(ctx->pass_run);
    ctx->pass_run = r = nullptr;
    std::vector<ofx_opo_spec> specs((size_t)batch);
    for (int b = 0; b < batch; b++) {
      ofx_opo_spec& s = specs[b];
      memset(&s, 0, sizeof s);
      s.pass = *p;
      if (!want_ap) s.pass.n_cols = 0;
      s.beta = beta.data();
      s.alpha = alpha.data();
      s.G = Gd.data() + (g_batch == 1 ? 0 : (size_t)b * p->nsteps);
      s.field = field.data() + 2 * n * b;
    }
    st = ofx_opo_create(ctx, prec, batch, specs.data(), 0, tune, &r);
    if (st) return st;
    ctx->pass_run = r;
    ctx->pass_key = key;
  }
  const size_t cs = r->csz;
  // the MEX contract returns the input ApFT/stepMods with written entries replaced
  for (int b = 0; b < batch && !st; b++) {
    if (want_ap) {
      std::vector<double> tmp(2 * n * p->n_cols);
      for (size_t j = 0; j < tmp.size(); j++) tmp[j] = (double)ApFT[2 * n * p->n_cols * b + j];
      st = upload(r, r->bufs[b].apft, convert(tmp.data(), tmp.size(), prec, keep), n * p->n_cols * cs);
    }
    if (!st && stepMods) {
      keep.emplace_back((size_t)p->nsteps * 4);
      float* f = (float*)keep.back().data();
      for (int j = 0; j < p->nsteps; j++) f[j] = (float)stepMods[(size_t)b * p->nsteps + j];
      st = upload(r, r->bufs[b].stepmods, f, (size_t)p->nsteps * 4);
    }
  }
  if (!st) st = ofx_opo_launch(r, 0, 0.0);
  int sync_st = st ? st : ofx_opo_sync(r);
  if (sync_st == OFX_E_HIP || sync_st == OFX_E_TIMEOUT || st) {
    // a run in an unknown state is not kept
    ofx_opo_destroy(ctx->pass_run);
    ctx->pass_run = nullptr;
    return st ? st : sync_st;
  }
  std::vector<double> f(2 * n), ap(2 * n * std::max(1, p->n_cols)), sm((size_t)p->nsteps);
  for (int b = 0; b < batch && (st == OFX_OK); b++) {
    int s_b = 0;
    st = ofx_opo_read(r, b, f.data(), nullptr, nullptr, &s_b, nullptr, nullptr,
                      want_ap ? ap.data() : nullptr, stepMods ? sm.data() : nullptr);
    if (st) break;
    for (size_t j = 0; j < 2 * n; j++) Et[2 * n * b + j] = (T)f[j];
    if (want_ap)
      for (size_t j = 0; j < 2 * n * p->n_cols; j++) ApFT[2 * n * p->n_cols * b + j] = (T)ap[j];
    if (stepMods)
      for (int j = 0; j < p->nsteps; j++) stepMods[(size_t)b * p->nsteps + j] = (T)sm[j];
    if (status) status[b] = s_b;
  }
  if (st) {
    ofx_opo_destroy(ctx->pass_run);
    ctx->pass_run = nullptr;
    return st;
  }
  return sync_st;
}

extern "C" {

int ofx_crystal_pass_host(ofx_context* ctx, const ofx_pass_params* p, int batch, float* Et,
                          const float* alpha_w, const float* G, int g_batch, const float* beta_op,
                          float* ApFT, float* stepMods, int* status, const ofx_tuning* tune) {
  return pass_host_impl<float>(ctx, OFX_F32, p, batch, Et, alpha_w, G, g_batch, beta_op, ApFT,
                               stepMods, status, tune);
}

int ofx_crystal_pass_host_f64(ofx_context* ctx, const ofx_pass_params* p, int batch, double* Et,
                              const double* alpha_w, const double* G, int g_batch,
                              const double* beta_op, double* ApFT, double* stepMods, int* status,
                              const ofx_tuning* tune) {
  return pass_host_impl<double>(ctx, OFX_F64, p, batch, Et, alpha_w, G, g_batch, beta_op, ApFT,
                                stepMods, status, tune);
}

}  // extern "C"

// ---------------- single pass without allocation ----------------
// The library's fixed device array (be_scratch) holds everything the pass
// needs besides the caller's arrays.  Layout per call, each block 256-byte
// aligned: the tag epoch pair (always first, so the tag base only grows
// from call to call), the per-instance state words, the descriptors, the
// clusters' barrier words and exchange buffers (cleared every call, so no
// word left by an earlier call of another shape can carry an awaited tag),
// then the tables.  Host staging lives in file-scope arrays; a mutex
// serializes the calls, which share both.
namespace {

constexpr int NA_MAX_BATCH = 16;
constexpr int NA_MAX_N = 65536;
constexpr int NA_MAX_GRID = 4096;
constexpr int NA_MAX_XCD = 64;

std::mutex g_na_mu;
double na_tw64[2 * NA_MAX_N], na_wt64[2 * NA_MAX_N], na_ab64[NA_MAX_N];
float na_tw32[2 * NA_MAX_N], na_wt32[2 * NA_MAX_N], na_ab32[NA_MAX_N];
InstDesc na_inst[NA_MAX_BATCH];
ClusterDesc na_clus[NA_MAX_BATCH];
int na_order[NA_MAX_BATCH], na_status[NA_MAX_BATCH];
int na_wgc[NA_MAX_GRID], na_wgr[NA_MAX_GRID];

size_t na_take(size_t& off, size_t bytes) {
  const size_t o = off;
  off = (off + bytes + 255) & ~(size_t)255;
  return o;
}

}  // namespace

extern "C" int ofx_crystal_pass_noalloc(ofx_context* ctx, const ofx_pass_params* p, int batch,
                                        float* Et, const float* alpha_w, const float* G,
                                        int g_batch, const float* beta_op, float* ApFT,
                                        float* stepMods, int* status) {
  if (!ctx) return OFX_E_ARG;
  int st = check_pass(ctx, p);
  if (st) return st;
  if (batch == 0) return OFX_OK;
  if (batch < 0 || !Et || !alpha_w || !G || !beta_op)
    return fail(ctx, OFX_E_ARG, "Et, alpha_w, G, beta_op are required and batch >= 0");
  if (g_batch != 1 && g_batch != batch)
    return fail(ctx, OFX_E_ARG, "G must hold NSteps or NSteps*batch values (g_batch=%d)", g_batch);
  if (batch > NA_MAX_BATCH)
    return fail(ctx, OFX_E_ARG, "the no-allocation pass takes at most %d instances (got %d)",
                NA_MAX_BATCH, batch);
  std::lock_guard<std::mutex> lock(g_na_mu);
  unsigned char* base = nullptr;
  size_t cap = 0;
  if (be_scratch(ctx->be, &base, &cap))
    return fail(ctx, OFX_E_UNSUPPORTED, "no device array for the no-allocation pass");
  const int n = p->n, nsteps = p->nsteps, method = p->method;
  const int n_cols = ApFT ? p->n_cols : 0;

  // geometry: as ofx_opo_create plans a batch of equal instances (one
  // cluster per instance, all resident at once)
  int C = 0, n1 = 0, block = 0;
  st = plan_instance(ctx, *p, batch, nullptr, 0, 0, 8, &C, &n1);
  if (!st) st = plan_block(ctx, nullptr, C, &block);
  if (st) return st;
  const int n2 = n / n1;
  const int xmode = exchange_mode(0, C);
  const size_t lds = lds_bytes_for(n1, n2, C, method, 8);
  const int bpc = be_blocks_per_cu(ctx->be, OFX_F32, xmode, block, lds);
  if (bpc < 1)
    return fail(ctx, OFX_E_ARG, "kernel cannot launch with block=%d and %zu bytes of LDS", block, lds);
  const int slots = ctx->info.cus * bpc;
  if ((long long)batch * C > slots || slots > NA_MAX_GRID)
    return fail(ctx, OFX_E_ARG, "%d instances of %d workgroups do not fit the device at once", batch,
                C);
  // XCD-aware placement: workgroup id g runs on XCD g % nx; a cluster goes
  // to the XCD with most free slots, spread only when none has room
  const int nx = std::max(1, std::min(ctx->info.nxcd, NA_MAX_XCD));
  const int per_x = slots / nx;
  int freex[NA_MAX_XCD];
  for (int x = 0; x < nx; x++) freex[x] = per_x;
  for (int g = 0; g < slots; g++) na_wgc[g] = -1, na_wgr[g] = 0;
  int maxid = -1;
  auto fullest = [&]() {
    int best = 0;
    for (int x = 1; x < nx; x++)
      if (freex[x] > freex[best]) best = x;
    return best;
  };
  for (int c = 0; c < batch; c++) {
    int x = fullest();
    const bool one = freex[x] >= C;
    for (int q = 0; q < C; q++) {
      if (!one) x = fullest();
      const int id = x + nx * (per_x - freex[x]);
      freex[x]--;
      na_wgc[id] = c;
      na_wgr[id] = q;
      maxid = std::max(maxid, id);
    }
  }
  const int grid = maxid + 1;

  // layout of the device array
  size_t off = 0;
  const size_t o_epoch = na_take(off, 8);
  const size_t o_clear = off;  // from here to o_tables: cleared every call
  const size_t o_state = na_take(off, 32 * (size_t)batch);
  const size_t o_inst = na_take(off, sizeof(InstDesc) * batch);
  const size_t o_clus = na_take(off, sizeof(ClusterDesc) * batch);
  const size_t o_order = na_take(off, 4 * (size_t)batch);
  const size_t o_wgc = na_take(off, 4 * (size_t)grid);
  const size_t o_wgr = na_take(off, 4 * (size_t)grid);
  const size_t o_xid = na_take(off, 4 * (size_t)grid);
  const size_t o_sync = na_take(off, 128 * (size_t)batch);
  size_t o_t1[NA_MAX_BATCH], o_t2[NA_MAX_BATCH], o_slots[NA_MAX_BATCH], o_xcc[NA_MAX_BATCH],
      o_kx[NA_MAX_BATCH];
  const size_t tsz = 16;  // one tagged complex64 sample
  for (int c = 0; c < batch; c++) {
    o_t1[c] = na_take(off, 2 * (size_t)n * tsz);
    o_t2[c] = na_take(off, 2 * (size_t)n * tsz);
    o_slots[c] = na_take(off, 2 * 8 * 4 * (size_t)C * tsz);
    o_xcc[c] = na_take(off, 4 * (size_t)C);
    o_kx[c] = method == OFX_METHOD_RK23 ? na_take(off, 2 * (size_t)C * 6 * (size_t)n2 * tsz) : 0;
  }
  const size_t o_tables = off;
  const size_t o_tw = na_take(off, 8 * (size_t)n);
  const size_t o_wt = na_take(off, 8 * (size_t)n);
  const size_t o_ab = na_take(off, 4 * (size_t)n);
  const size_t o_beta = na_take(off, 8 * (size_t)n);
  const size_t o_alpha = na_take(off, 8 * (size_t)n);
  const size_t o_G = na_take(off, 8 * (size_t)nsteps * (size_t)g_batch);
  if (off > cap)
    return fail(ctx, OFX_E_ARG,
                "N=%d, NSteps=%d, batch=%d needs %zu bytes of device scratch; the no-allocation "
                "pass holds %zu",
                n, nsteps, batch, off, cap);

  // tables on the host (the same arithmetic as ofx_opo_create), FP32
  fill_twiddle(n, na_tw64);
  time_tables(n, p->dt, p->w0, na_wt64, na_ab64);
  for (int j = 0; j < 2 * n; j++) {
    na_tw32[j] = (float)na_tw64[j];
    na_wt32[j] = (float)na_wt64[j];
  }
  for (int j = 0; j < n; j++) na_ab32[j] = (float)na_ab64[j];

  unsigned char* const S = base;
  int* const st_d = (int*)(S + o_state);
  for (int b = 0; b < batch; b++) {
    InstDesc& d = na_inst[b];
    memset(&d, 0, sizeof d);
    d.n = n;
    d.n1 = n1;
    d.n2 = n2;
    d.lg1 = ilog2i(n1);
    d.lg2 = ilog2i(n2);
    d.nsteps = nsteps;
    d.n_cols = n_cols;
    d.method = method;
    d.w0 = p->w0;
    d.bdiffw0 = p->bdiffw0;
    d.h = p->h;
    d.dt = p->dt;
    d.max_err = p->max_err;
    d.min_err = p->min_err;
    d.sel = p->sel;
    d.beta = (const double*)(S + o_beta);
    d.alpha = (const double*)(S + o_alpha);
    d.G = (const double*)(S + o_G) + (g_batch == 1 ? 0 : (size_t)b * nsteps);
    d.wt = S + o_wt;
    d.absorb = S + o_ab;
    d.field = Et + 2 * (size_t)n * b;
    d.apft = n_cols ? (void*)(ApFT + 2 * (size_t)n * n_cols * b) : nullptr;
    d.stepmods = stepMods ? stepMods + (size_t)nsteps * b : nullptr;
    d.status = st_d + b;
    d.trips_done = st_d + batch + b;
    d.fresh = st_d + 2 * batch + b;
    d.counters = (long long*)(S + o_state + 16 * (size_t)batch) + 2 * b;
    ClusterDesc& cd = na_clus[b];
    memset(&cd, 0, sizeof cd);
    cd.C = C;
    cd.first_inst = b;
    cd.n_inst = 1;
    cd.halo = halo_of(method);
    cd.t1 = S + o_t1[b];
    cd.t2 = S + o_t2[b];
    cd.slots = S + o_slots[b];
    cd.bar = (unsigned*)(S + o_sync + 128 * (size_t)b);
    cd.abort_flag = (int*)(S + o_sync + 128 * (size_t)b + 64);
    cd.xcc = (int*)(S + o_xcc[b]);
    cd.kx = method == OFX_METHOD_RK23 ? S + o_kx[b] : nullptr;
    cd.epoch = (unsigned*)(S + o_epoch);  // one tag base for every cluster
    na_order[b] = b;
  }

  Backend* be = ctx->be;
  auto hip_err = [&](const char* what) {
    return fail(ctx, OFX_E_HIP, "%s failed: %s", what, be_error(be));
  };
  if (be_fill_zero(be, S + o_clear, o_tables - o_clear)) return hip_err("clearing the scratch");
  if (be_h2d(be, S + o_inst, na_inst, sizeof(InstDesc) * batch) ||
      be_h2d(be, S + o_clus, na_clus, sizeof(ClusterDesc) * batch) ||
      be_h2d(be, S + o_order, na_order, 4 * (size_t)batch) ||
      be_h2d(be, S + o_wgc, na_wgc, 4 * (size_t)grid) ||
      be_h2d(be, S + o_wgr, na_wgr, 4 * (size_t)grid) ||
      be_h2d(be, S + o_tw, na_tw32, 8 * (size_t)n) ||
      be_h2d(be, S + o_wt, na_wt32, 8 * (size_t)n) ||
      be_h2d(be, S + o_ab, na_ab32, 4 * (size_t)n))
    return hip_err("upload");
  if (be_widen(be, (double*)(S + o_beta), beta_op, (size_t)n) ||
      be_widen(be, (double*)(S + o_alpha), alpha_w, (size_t)n) ||
      be_widen(be, (double*)(S + o_G), G, (size_t)nsteps * g_batch))
    return hip_err("table conversion");
  LaunchArgs a;
  memset(&a, 0, sizeof a);
  a.inst = (const InstDesc*)(S + o_inst);
  a.clusters = (const ClusterDesc*)(S + o_clus);
  a.inst_order = (const int*)(S + o_order);
  a.wg_cluster = (const int*)(S + o_wgc);
  a.wg_rank = (const int*)(S + o_wgr);
  a.twiddle = S + o_tw;
  a.nmax = n;
  a.trips = 0;
  a.conv_tol = 0.0;
  a.timeout_ticks = (unsigned long long)((double)ctx->info.wallclock_khz * 1000.0 * 10.0);
  a.xcc_ids = (int*)(S + o_xid);
  if (be_launch(be, a, OFX_F32, xmode, grid, block, lds, nullptr, nullptr, 0,
                (unsigned*)(S + o_epoch), 1, nullptr, 0))
    return hip_err("launch");
  if (status && be_d2d(be, status, st_d, 4 * (size_t)batch)) return hip_err("status copy");
  if (be_d2h(be, na_status, st_d, 4 * (size_t)batch) || be_sync(be)) return hip_err("pass");
  for (int b = 0; b < batch; b++)
    if (na_status[b] == ST_BARRIER_TIMEOUT)
      return fail(ctx, OFX_E_TIMEOUT, "instance %d: a cluster member timed out waiting for exchange data", b);
  for (int b = 0; b < batch; b++)
    if (na_status[b] == ST_NONFINITE)
      return fail(ctx, OFX_E_NONFINITE, "instance %d: field became non-finite", b);
  return OFX_OK;
}
