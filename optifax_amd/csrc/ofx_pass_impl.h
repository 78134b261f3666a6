// Body of the persistent cluster kernel that runs adaptive split-step crystal
// passes (and, in round-trip mode, whole OPO round trips) for a batch of
// instances.
//
// Work decomposition (one instance, N = n1 * n2 points, C workgroups):
//   time domain   : point j = c + n1*r. Workgroup `rank` owns the W = n1/C
//                   columns c in [rank*W, rank*W+W) plus `halo` columns on each
//                   side (halo copies are recomputed redundantly, so the
//                   nonlinear stages need no communication).
//   frequency     : bin k = k2 + n2*k1. Workgroup `rank` owns the R = n2/C
//                   rows k2 in [rank*R, rank*R+R), all k1.
//   one FFT       : four-step (Bailey 1990): length-n2 transforms of the own
//                   columns in LDS, twiddle, global transpose (buffer t1),
//                   length-n1 transforms of the own rows in LDS.  The inverse
//                   mirrors it through buffer t2.  Each transform is a radix-4
//                   Stockham autosort FFT (with one radix-2 pass for odd log2).
//   one accepted step = 2 cluster barriers (t1 and t2 transposes); the error
//                   max-norm travels with the t1 barrier, so a rejected step
//                   costs 1 barrier.  Step control is replicated in every
//                   workgroup from identical inputs, so no broadcast is needed.
//
// This header compiles as HIP device code (ofx_kernels.hip) and, with OFX_EMU
// defined, as plain C++ where each workgroup is one CPU thread; the emulation
// build exists only to test indexing and control flow on a machine without a
// GPU.
#pragma once
#include "ofx_internal.h"
#include <math.h>
#include <float.h>

#ifdef OFX_EMU
#include <sched.h>

#include <cmath>
#define OFX_DEV inline
#define OFX_TID 0
#define OFX_NT 1
#define OFX_SYNC() ((void)0)
namespace ofx {
bool emu_cluster_barrier(const ClusterDesc* cd, unsigned target, unsigned long long timeout);
unsigned long long emu_clock();  // wall clock in the device's ticks
}
#else
#define OFX_DEV __device__ __forceinline__
#define OFX_TID ((int)threadIdx.x)
#define OFX_NT ((int)blockDim.x)
#define OFX_SYNC() __syncthreads()
#endif

namespace ofx {

// Hardware register holding the accelerator-complex-die id of the executing
// wave on gfx940-class parts (HW_REG_XCC_ID, bits [3:0]).
#define OFX_HWREG_XCC_ID ((20) | (0 << 6) | ((4 - 1) << 11))

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
OFX_DEV cpx<R> mkc(R a, R b) {
  cpx<R> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename R>
OFX_DEV cpx<R> cadd(cpx<R> a, cpx<R> b) { return mkc<R>(a.x + b.x, a.y + b.y); }
template <typename R>
OFX_DEV cpx<R> csub(cpx<R> a, cpx<R> b) { return mkc<R>(a.x - b.x, a.y - b.y); }
template <typename R>
OFX_DEV cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
OFX_DEV cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return mkc<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
OFX_DEV cpx<R> cscale(cpx<R> a, R s) { return mkc<R>(a.x * s, a.y * s); }

// Global-memory views of device buffers held in plain pointer types: with
// them the compiler emits global instead of flat accesses (a flat access
// also counts as an LDS access, so every LDS wait after it waits for the
// memory round trip too).
#ifdef OFX_EMU
#define OFX_GLOBAL
#else
#define OFX_GLOBAL __attribute__((address_space(1)))
#endif
template <typename T>
OFX_DEV const OFX_GLOBAL T* gl(const T* p) {
  return (const OFX_GLOBAL T*)p;
}
template <typename T>
OFX_DEV OFX_GLOBAL T* gl(T* p) {
  return (OFX_GLOBAL T*)p;
}
template <typename R, typename I>
OFX_DEV cpx<R> ldg(const OFX_GLOBAL cpx<R>* p, I i) {
  return mkc<R>(p[i].x, p[i].y);
}
template <typename R, typename I>
OFX_DEV void stg(OFX_GLOBAL cpx<R>* p, I i, cpx<R> v) {
  p[i].x = v.x;
  p[i].y = v.y;
}

// A 4- or 8-byte value that is the same in every lane, marked as such (the
// compiler then keeps it in scalar registers instead of one per lane)
template <typename T>
OFX_DEV T uniform(T v) {
#ifdef OFX_EMU
  return v;
#else
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values only");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
  } else {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned long long hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(T, (hi << 32) | lo);
  }
#endif
}

// x - 2*pi*round(x/(2*pi)) with a two-term (Cody-Waite) split of 2*pi; the
// fma makes n*TWO_PI_HI exact, so the reduction is accurate for |x| < 2^40.
OFX_DEV double reduce_2pi(double x) {
  const double TWO_PI_HI = 6.283185307179586232;
  const double TWO_PI_LO = 2.4492935982947064e-16;
  const double INV_TWO_PI = 0.15915494309189533577;
  double q = rint(x * INV_TWO_PI);
  double r = fma(-q, TWO_PI_HI, x);
  return fma(-q, TWO_PI_LO, r);
}

// exp(i*x) for an FP64 argument: reduced in FP64, evaluated in R.
// FP32: x = q pi/2 + r, |r| <= pi/4, with q = rint(2x/pi) and r from a
// two-part pi/2 in FP64 (Cody & Waite's reduction; the fma keeps q * hi
// exact), then the Taylor polynomials of sin r (to r^9) and cos r (to r^10)
// in FP32, whose truncation stays below 2e-9 on |r| <= pi/4, rotated by q's
// quadrant.  Largest error of exp(ix) over |x| <= 1e10: 9.4e-8 (FP32 ulp(1)
// is 1.2e-7).
OFX_DEV cpx<float> expi(double x, float) {
  const double PIO2_HI = 1.5707963267948966, PIO2_LO = 6.123233995736766e-17;
  const double q = rint(x * 0.63661977236758134);
  const float r = (float)fma(-q, PIO2_LO, fma(-q, PIO2_HI, x));
  const float r2 = r * r;
  const float s =
      r * fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f),
                        -1.0f / 6.0f),
               1.0f);
  const float c = fmaf(
      r2,
      fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f), -1.0f / 720.0f),
                    1.0f / 24.0f),
           -0.5f),
      1.0f);
  const int k = (int)fma(-4.0, floor(0.25 * q), q);   // q mod 4 (exact): exp(i q pi/2) = i^k
  const float re = k == 0 ? c : (k == 1 ? -s : (k == 2 ? -c : s));
  const float im = k == 0 ? s : (k == 1 ? c : (k == 2 ? -s : -c));
  return mkc<float>(re, im);
}
OFX_DEV cpx<double> expi(double x, double) {
  double r = reduce_2pi(x);
#ifdef OFX_EMU
  return mkc<double>(cos(r), sin(r));
#else
  double s, c;
  sincos(r, &s, &c);
  return mkc<double>(c, s);
#endif
}

// Nonlinear polarisation term of the NEE (NEE_CPU.m:98-101):
//   nl = A^2 e^{i th} + 2 |A|^2 e^{-i th}
template <typename R>
OFX_DEV cpx<R> nl_term(cpx<R> A, double th) {
  cpx<R> e = expi(th, R(0));
  R a2 = A.x * A.x + A.y * A.y;
  R sx = A.x * A.x - A.y * A.y;
  R sy = R(2) * A.x * A.y;
  return mkc<R>(sx * e.x - sy * e.y + R(2) * a2 * e.x, sx * e.y + sy * e.x - R(2) * a2 * e.y);
}

// same with the unit phasor e = e^{i th} given
template <typename R>
OFX_DEV cpx<R> nl_term_e(cpx<R> A, cpx<R> e) {
  R a2 = A.x * A.x + A.y * A.y;
  R sx = A.x * A.x - A.y * A.y;
  R sy = R(2) * A.x * A.y;
  return mkc<R>(sx * e.x - sy * e.y + R(2) * a2 * e.x, sx * e.y + sy * e.x - R(2) * a2 * e.y);
}

// K = -i*g*(nl - i*cc*(up - dn)),  g = h*s*G (NEE_CPU.m:103-105)
template <typename R>
OFX_DEV cpx<R> k_term(cpx<R> nl, cpx<R> up, cpx<R> dn, R g, R cc) {
  R dx = up.x - dn.x, dy = up.y - dn.y;
  R ix = nl.x + cc * dy;
  R iy = nl.y - cc * dx;
  return mkc<R>(g * iy, -g * ix);
}

// ------------------------------------------------------------------------
// Batched Stockham FFT in LDS.  nseq sequences of length L = 2^lg, element e
// of sequence q at buf[q*ss + e]: radix-4 passes (one radix-2 pass for odd
// lg).  Ping-pongs between `in` and `out`; returns the buffer that holds the
// result.  SIGN = -1 forward (unnormalised), +1 inverse (unnormalised).
// tw: the passes' twiddle tables (pass_twiddle_index).  (Measured on the
// ring workload: radix-8 passes with register DFTs were 6 % slower per step.)
// ------------------------------------------------------------------------
// Twiddle tables of the radix-4 passes of a length-2^lg transform: the pass
// with stride Ns > 1 reads w_c[k] = exp(-2 pi i c k / (4 Ns)), c = 1, 2, 3,
// k < Ns, stored as three arrays of Ns entries from offset Ns - Ns0 (Ns0 = 2
// for odd lg, 4 for even lg), so that consecutive butterflies read
// consecutive entries (L - Ns0 entries in all).  pass_twiddle_index gives,
// for entry j, the index into exp(-2 pi i m / L) it holds.
OFX_DEV int pass_ns0(int lg) { return (lg & 1) ? 2 : 4; }
OFX_DEV int pass_twiddle_index(int lg, int j) {
  int ns = pass_ns0(lg), off = 0;
  while (j >= off + 3 * ns) {
    off += 3 * ns;
    ns <<= 2;
  }
  const int c = (j - off) / ns + 1, k = (j - off) - (c - 1) * ns;
  return c * k * ((1 << lg) / (4 * ns));
}

template <typename R, int SIGN>
OFX_DEV cpx<R>* lds_fft(cpx<R>* in, cpx<R>* out, int lg, int nseq, int ss, const cpx<R>* tw) {
  typedef cpx<R> C;
  const int L = 1 << lg;
  const int tid = OFX_TID, nt = OFX_NT;
  int Ns = 1;
  if (lg & 1) {
    const int h = L >> 1;
    const int nb = nseq << (lg - 1);
    for (int t = tid; t < nb; t += nt) {
      const int q = t >> (lg - 1), j = t & (h - 1);
      const C* ib = in + q * ss;
      C* ob = out + q * ss;
      C v0 = ib[j], v1 = ib[j + h];
      ob[2 * j] = cadd(v0, v1);
      ob[2 * j + 1] = csub(v0, v1);
    }
    OFX_SYNC();
    C* tmp = in;
    in = out;
    out = tmp;
    Ns = 2;
  }
  if (lg >= 2) {
    const int q4 = L >> 2;
    const int nb = nseq << (lg - 2);
    const int ns0 = pass_ns0(lg);
    for (; Ns < L; Ns <<= 2) {
      const C* tp = tw + (Ns - ns0);   // this pass's w_1, w_2, w_3 (Ns > 1)
      for (int t = tid; t < nb; t += nt) {
        const int q = t >> (lg - 2), j = t & (q4 - 1);
        const int k = j & (Ns - 1);
        const C* ib = in + q * ss;
        C* ob = out + q * ss;
        C v0 = ib[j], v1 = ib[j + q4], v2 = ib[j + 2 * q4], v3 = ib[j + 3 * q4];
        if (Ns > 1) {
          C w1 = tp[k], w2 = tp[Ns + k], w3 = tp[2 * Ns + k];
          if (SIGN > 0) {
            v1 = cmulc(v1, w1);
            v2 = cmulc(v2, w2);
            v3 = cmulc(v3, w3);
          } else {
            v1 = cmul(v1, w1);
            v2 = cmul(v2, w2);
            v3 = cmul(v3, w3);
          }
        }
        C s02 = cadd(v0, v2), d02 = csub(v0, v2), s13 = cadd(v1, v3), d13 = csub(v1, v3);
        // forward multiplies d13 by -i, inverse by +i
        C rd = (SIGN < 0) ? mkc<R>(d13.y, -d13.x) : mkc<R>(-d13.y, d13.x);
        const int base = ((j - k) << 2) + k;
        ob[base] = cadd(s02, s13);
        ob[base + Ns] = cadd(d02, rd);
        ob[base + 2 * Ns] = csub(s02, s13);
        ob[base + 3 * Ns] = csub(d02, rd);
      }
      OFX_SYNC();
      C* tmp = in;
      in = out;
      out = tmp;
    }
  }
  return in;
}

// ------------------------------------------------------------------------
// Cluster exchange.
//
// Workgroups of a cluster exchange data only through tagged buffers: the
// transposes t1/t2, the per-wave reduction records (slots) and, for RK23, the
// boundary columns of K4 (kx).  Every 32-bit word of an exchanged value
// travels in a 64-bit word whose upper half is the exchange's tag; the words
// are written and read with device-scope relaxed 64-bit atomics (sc1
// accesses, which bypass the non-coherent caches, so no agent-scope fence and
// no L2 writeback is involved).  A reader polls the words it needs until each
// carries the tag it expects: the data is its own completion flag.  An
// exchange therefore costs the stores' trip to the coherent level plus the
// reader's poll, with no store-completion wait, counter atomic or flag read
// in between.  That wins when each thread reads one or two samples per
// exchange (clusters of 64 workgroups and more).  With larger shares the
// doubled bytes and the repeated polls cost more, and the kernel's other
// instantiation (XM_BARRIER) stores the samples untagged (plain form, half
// the bytes) and meets at a counter barrier before reading them once.
//
// Tags: tag = base + x, x = exchanges so far in this launch (every member
// runs the same sequence of exchanges), base = the cluster's epoch word,
// moved before each launch past every tag the previous launch used.  Every
// buffer is double-buffered by the parity of x.  A member rewrites a parity
// half two exchanges after its previous use, and only after it has read data
// that every member produced after finishing its reads of that use, so a
// word still in a buffer always carries an older tag than the one awaited.
// A poll gives up, and aborts the cluster, after the launch's time limit or
// when another member has aborted.
//
// Cluster-wide reductions go through the slots: every wave reduces its lanes
// and publishes one record (no workgroup reduction before the exchange); the
// readers reduce the C * waves records, one per thread.  op 0 = max (one
// value), op 1 = sum (three values, combined in a fixed order, so all members
// obtain bit-identical results: the step-control decisions depend on it).
// Slot layout: [parity][wave < 8][4 values][rank], each value a tagged
// element the size of one complex sample.
// ------------------------------------------------------------------------
#define OFX_SLOT_WAVES 8
typedef unsigned long long u64;

OFX_DEV double red_op(double a, double b, int op) { return op ? a + b : (a > b ? a : b); }

// (global address space on the device: global_* rather than flat_*
// instructions, which would also wait on the LDS counter)
OFX_DEV void tag_store(u64* p, u64 v) {
#ifdef OFX_EMU
  __atomic_store_n(p, v, __ATOMIC_RELAXED);
#else
  __hip_atomic_store((__attribute__((address_space(1))) u64*)p, v, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
#endif
}

OFX_DEV u64 tag_load(const u64* p) {
#ifdef OFX_EMU
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#else
  return __hip_atomic_load((__attribute__((address_space(1))) u64*)const_cast<u64*>(p),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}

OFX_DEV int flag_load(const int* p) {
#ifdef OFX_EMU
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#else
  return __hip_atomic_load((__attribute__((address_space(1))) int*)const_cast<int*>(p),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}

OFX_DEV void flag_set(int* p) {
#ifdef OFX_EMU
  __atomic_store_n(p, 1, __ATOMIC_RELAXED);
#else
  __hip_atomic_store((__attribute__((address_space(1))) int*)p, 1, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
#endif
}

// A value of type T (a multiple of 4 bytes) as NW tagged 64-bit words.
template <typename T>
struct Tagged {
  static constexpr int NW = (int)(sizeof(T) / 4);
  struct Words {
    unsigned w[NW];
  };
  // Tagged form: word i of the element at p lives at p + i * ws.  Plain form
  // (counter-barrier exchanges): the value's own bytes, NW / 2 words at p;
  // `load` then supplies the expected tag itself.
  static OFX_DEV void put(u64* dst, size_t ws, int plain, T v, unsigned tag) {
    const Words x = __builtin_bit_cast(Words, v);
    if (plain) {
#pragma unroll
      for (int i = 0; i < NW / 2; i++)
        tag_store(dst + i, (u64)x.w[2 * i] | ((u64)x.w[2 * i + 1] << 32));
      return;
    }
    const u64 hi = (u64)tag << 32;
#pragma unroll
    for (int i = 0; i < NW; i++) tag_store(dst + i * ws, hi | (u64)x.w[i]);
  }
  static OFX_DEV void load(const u64* src, size_t ws, int plain, unsigned tag, u64* w) {
    if (plain) {
      const u64 hi = (u64)tag << 32;
#pragma unroll
      for (int i = 0; i < NW / 2; i++) {
        const u64 q = tag_load(src + i);
        w[2 * i] = hi | (q & 0xffffffffull);
        w[2 * i + 1] = hi | (q >> 32);
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < NW; i++) w[i] = tag_load(src + i * ws);
  }
  static OFX_DEV bool check(const u64* w, unsigned tag) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NW; i++) ok = ok & ((unsigned)(w[i] >> 32) == tag);
    return ok;
  }
  static OFX_DEV T value(const u64* w) {
    Words x;
#pragma unroll
    for (int i = 0; i < NW; i++) x.w[i] = (unsigned)w[i];
    return __builtin_bit_cast(T, x);
  }
};

// Where a tagged element lives: its first word and the distance between its
// words.  Buffers of complex samples keep word i of all elements in plane i
// (a plane of P words per parity half), so the stores and loads of one word
// index cover contiguous memory across the lanes.  In plain form (clusters
// that meet at a counter barrier before reading) element e is stored
// untagged at e * NW / 2: half the bytes, which is what bounds the exchange
// when each workgroup holds a large share of the field.
struct TagPtr {
  const u64* p;
  size_t ws;
  int plain;
};
template <typename T>
OFX_DEV TagPtr tp_plane(int plain, const u64* half, size_t P, size_t e) {
  if (plain) return TagPtr{half + e * (Tagged<T>::NW / 2), 1, 1};
  return TagPtr{half + e, P, 0};
}

// Poll state of one wait: the limit counts from the first unsuccessful poll.
struct Spin {
  int* abort_flag;
  unsigned long long limit, t0;
  int n;
};

OFX_DEV Spin spin_begin(const ClusterDesc& cd, unsigned long long limit) {
  Spin s;
  s.abort_flag = cd.abort_flag;
  s.limit = limit;
  s.t0 = 0;
  s.n = 0;
  return s;
}

// Pause before the next poll round; false when the cluster is to give up
// (the clock is read every round, the abort flag every 8th round from the
// 8th on: most waits end within a few rounds, and a flag read in the first
// one would add its own round trip to every wait).
OFX_DEV bool spin_wait(Spin& s) {
#ifdef OFX_EMU
  sched_yield();
  const unsigned long long t = emu_clock();
  if (s.n++ == 0) s.t0 = t;
  if (flag_load(s.abort_flag)) return false;
  if (t - s.t0 > s.limit) {
    flag_set(s.abort_flag);
    return false;
  }
  return true;
#else
  __builtin_amdgcn_s_sleep(1);
  const unsigned long long t = wall_clock64();
  if (s.n == 0) s.t0 = t;
  if ((++s.n & 7) == 0 && flag_load(s.abort_flag)) return false;
  if (t - s.t0 > s.limit) {
    flag_set(s.abort_flag);
    return false;
  }
  return true;
#endif
}

// Poll `cnt` tagged elements of type T held by this thread: element j is at
// src(j) (a TagPtr) and goes to commit(j, value) once all its words carry `tag`.
// Up to J elements (four words) are in flight at once (more costs registers
// that the rest of the kernel needs).  False when the cluster gave up.
template <typename T, typename Src, typename Commit>
OFX_DEV bool poll_tagged(int cnt, unsigned tag, Src src, Commit commit, Spin& sp) {
  typedef Tagged<T> TT;
  constexpr int J = 4 / TT::NW;
  for (int b = 0; b < cnt; b += J) {
    u64 w[J][TT::NW];
    unsigned pend = 0;
#pragma unroll
    for (int j = 0; j < J; j++)
      if (b + j < cnt) {
        const TagPtr t = src(b + j);
        TT::load(t.p, t.ws, t.plain, tag, w[j]);
        pend |= 1u << j;
      }
    for (;;) {
#pragma unroll
      for (int j = 0; j < J; j++)
        if (((pend >> j) & 1u) && TT::check(w[j], tag)) {
          commit(b + j, TT::value(w[j]));
          pend &= ~(1u << j);
        }
      if (!pend) break;
      if (!spin_wait(sp)) return false;
#pragma unroll
      for (int j = 0; j < J; j++)
        if ((pend >> j) & 1u) {
          const TagPtr t = src(b + j);
          TT::load(t.p, t.ws, t.plain, tag, w[j]);
        }
    }
  }
  return true;
}

// The same over a workgroup-wide index space [0, tot): this thread takes
// idx = OFX_TID + j * OFX_NT.  After a counter barrier (`complete`) the data
// are known to be in place and are read once, without tag checks.
template <typename T, typename Src, typename Commit>
OFX_DEV bool read_tagged(int tot, unsigned tag, Src src, Commit commit, Spin& sp, int complete) {
  if (complete) {
    typedef Tagged<T> TT;
    for (int idx = OFX_TID; idx < tot; idx += OFX_NT) {
      const TagPtr t = src(idx);
      u64 w[TT::NW];
      TT::load(t.p, t.ws, t.plain, tag, w);
      commit(idx, TT::value(w));
    }
    return true;
  }
  const int cnt = tot > OFX_TID ? (tot - OFX_TID + OFX_NT - 1) / OFX_NT : 0;
  return poll_tagged<T>(
      cnt, tag, [&](int j) { return src(OFX_TID + j * OFX_NT); },
      [&](int j, T v) { commit(OFX_TID + j * OFX_NT, v); }, sp);
}

// a double carried in a complex-sized tagged element
template <typename R>
struct Dbl {
  double d[sizeof(cpx<R>) / 8];
};
template <typename R>
OFX_DEV cpx<R> pack_d(double v) {
  Dbl<R> x;
  for (int i = 0; i < (int)(sizeof(cpx<R>) / 8); i++) x.d[i] = 0.0;
  x.d[0] = v;
  return __builtin_bit_cast(cpx<R>, x);
}
template <typename R>
OFX_DEV double unpack_d(cpx<R> c) {
  return __builtin_bit_cast(Dbl<R>, c).d[0];
}

OFX_DEV int slot_waves() {
#ifdef OFX_EMU
  return 1;
#else
  return (int)(blockDim.x >> 6);
#endif
}

// Workgroup-wide AND of the per-thread read verdicts.  After a counter
// barrier the reads cannot fail, and a plain barrier suffices.
OFX_DEV bool wg_all(bool ok, int complete) {
#ifdef OFX_EMU
  (void)complete;
  return ok;
#else
  if (complete) {
    __syncthreads();
    return true;
  }
  return __syncthreads_and(ok) != 0;
#endif
}

// first word of slot value i of (parity, rank, wave)
template <typename R>
OFX_DEV u64* slot_at(const ClusterDesc& cd, int par, int rank, int w, int i) {
  // [parity][wave][value][rank]: one value of every member is a contiguous run
  return (u64*)cd.slots +
         ((((size_t)par * OFX_SLOT_WAVES + w) * 4 + i) * cd.C + rank) * Tagged<cpx<R>>::NW;
}

template <typename R>
OFX_DEV void publish_slots(const ClusterDesc& cd, int par, int rank, int op, double a, double b,
                           double c, unsigned tag) {
  // op 0 (max) carries one value (the error quotient, +inf for a non-finite
  // field); op 1 (sums) carries three
  double v[3] = {a, b, c};
  const int nv = op ? 3 : 1;
#ifdef OFX_EMU
  const int w = 0;
#else
  for (int i = 0; i < nv; i++)
    for (int o = 32; o > 0; o >>= 1) v[i] = red_op(v[i], __shfl_xor(v[i], o), op);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) != 0) return;
#endif
  for (int i = 0; i < nv; i++)
    Tagged<cpx<R>>::put(slot_at<R>(cd, par, rank, w, i), 1, 0, pack_d<R>(v[i]), tag);
}

// Cluster sums of the three op-1 values.  Each thread adds its records in
// index order once each is complete; false when the cluster gave up.
template <typename R>
OFX_DEV bool gather_sums(const ClusterDesc& cd, int par, unsigned tag, double* out, double* red,
                         unsigned long long limit) {
  Spin sp = spin_begin(cd, limit);
  const int nw = slot_waves();
  const int tot = cd.C * nw;
  double v[3] = {0.0, 0.0, 0.0};
  bool ok = true;
  for (int t = OFX_TID; t < tot && ok; t += OFX_NT) {
    const int q = t / nw, w = t - q * nw;
    double r[3];
    ok = poll_tagged<cpx<R>>(
        3, tag, [&](int i) { return TagPtr{slot_at<R>(cd, par, q, w, i), 1, 0}; },
        [&](int i, cpx<R> x) { r[i] = unpack_d<R>(x); }, sp);
    if (ok)
      for (int i = 0; i < 3; i++) v[i] += r[i];
  }
#ifdef OFX_EMU
  (void)red;
  for (int i = 0; i < 3; i++) out[i] = v[i];
  return ok;
#else
  for (int i = 0; i < 3; i++)
    for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < 3; i++) red[3 * (threadIdx.x >> 6) + i] = v[i];
  ok = __syncthreads_and(ok) != 0;
  for (int i = 0; i < 3; i++) {
    double r = red[i];
    for (int w = 1; w < nw; w++) r += red[3 * w + i];
    out[i] = r;
  }
  return ok;
#endif
}

// Counter barrier (used once per instance, before the field is stored):
// every wave's stores have completed, then one thread bumps the cluster's
// counter and polls it up to `target`.  False when the cluster was aborted.
OFX_DEV bool cluster_barrier(const ClusterDesc& cd, unsigned target, unsigned long long timeout,
                             int* sflag) {
#ifdef OFX_EMU
  (void)sflag;
  return emu_cluster_barrier(&cd, target, timeout);
#else
  // the memory clobber keeps the compiler from moving any access across it
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    if (cd.C > 1) {
      __hip_atomic_fetch_add(cd.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long t0 = wall_clock64();
      unsigned polls = 0;
      // (the abort flag is read every 8th poll: a read per poll would double
      // the poll period)
      while (__hip_atomic_load(cd.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((++polls & 7) == 0 && flag_load(cd.abort_flag)) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > timeout) {
          flag_set(cd.abort_flag);
          ok = 0;
          break;
        }
      }
    }
    sflag[0] = ok;
  }
  __syncthreads();
  return sflag[0] != 0;
#endif
}

// Exchange counters of one cluster member.
struct SyncState {
  unsigned base;   // the cluster's tag base for this launch
  unsigned nx;     // exchanges so far
  unsigned nreal;  // counter barriers so far
  int* sflag;      // LDS word for the barrier's verdict
};

// Exchange modes (a template parameter of the kernel, so that each mode's
// code is compiled and register-allocated on its own).
#define XM_TAGS 0     // tagged words, readers poll the tags
#define XM_BARRIER 1  // plain words, a counter barrier before the reads

// Before the reads of an exchange.  In barrier mode the whole cluster meets
// first and the reads then find complete data; in tag mode the reads poll
// straight away.  False when the cluster gave up.
template <int XM>
OFX_DEV bool xchg_ready(const ClusterDesc& cd, SyncState& sy, unsigned long long limit) {
  if (XM == XM_TAGS) return true;
  sy.nreal++;
  return cluster_barrier(cd, sy.nreal * (unsigned)cd.C, limit, sy.sflag);
}

// ------------------------------------------------------------------------
// Per-instance state and geometry held by one workgroup.
// ------------------------------------------------------------------------
template <typename R>
struct Geo {
  int n, n1, n2, lg1, lg2, lgn;
  int Cn, lgC, rank, W, lgW, Rr, lgR, HW, LC, sT, sF, c0, k20;
  int plain;  // exchange in plain form after a counter barrier (else tagged, polled)
  cpx<R>*B0, *B1, *B2, *B3, *B4;  // B3, B4: RK23 only
  const cpx<R>* tw1;  // LDS: the length-n1 passes' twiddles (pass_twiddle_index), n1 entries
  const cpx<R>* tw2;  // LDS: the same for length n2
  const cpx<double>* twq;  // LDS, FP64: exp(-2 pi i q / n2), q < n2
  const cpx<double>* twr;  // LDS, FP64: exp(-2 pi i r / n), r < n1
  const cpx<R>* wl;   // LDS: exp(i w0 t_l), l < n1
  const cpx<R>* wh;   // LDS: exp(i w0 t_{h n1}), h < n2
  const R* ab;        // LDS: absorber factor of own column lc, row r at (lc - HW) n2 + r
};

// Four-step twiddle exp(-2 pi i m / n), m = q n1 + r < n, as the FP64
// product of the LDS entries exp(-2 pi i q / n2) and exp(-2 pi i r / n),
// rounded once to R (no access to an n-entry table in global memory; a
// product of two FP32 entries raised the FP32 pass error by half).
template <typename R>
OFX_DEV cpx<R> fourstep_tw(const Geo<R>& g, int m) {
  const cpx<double> a = g.twq[m >> g.lg1], b = g.twr[m & (g.n1 - 1)];
  return mkc<R>((R)(a.x * b.x - a.y * b.y), (R)(a.x * b.y + a.y * b.x));
}

template <typename R>
OFX_DEV int gidx(const Geo<R>& g, int lc, int r) {
  return ((g.c0 - g.HW + lc) + g.n1 * r) & (g.n - 1);
}

// exp(i w0 t_j) at local column lc, row r, from the two LDS factors
template <typename R>
OFX_DEV cpx<R> wt_at(const Geo<R>& g, int lc, int r) {
  const int j = gidx(g, lc, r);
  return cmul(g.wh[j >> g.lg1], g.wl[j & (g.n1 - 1)]);
}

// relative angular frequency of FFT-order bin k
OFX_DEV double dw_of(int k, int n, double dt) {
  const double TWO_PI = 6.283185307179586232;
  return TWO_PI * (double)(k < n / 2 ? k : k - n) / ((double)n * dt);
}

// Entry of sample (column c, row k2) in the transpose buffer t1.  Plain
// exchanges: row-major k2*n1 + c.  Tagged exchanges: member-blocked,
// ((c / W) n2 + k2) W + c % W, so a member's columns are one contiguous run
// for the writer and a row's reader takes runs of W samples.
template <typename R>
OFX_DEV size_t t1_index(const Geo<R>& g, int c, int k2) {
  if (g.plain) return (size_t)k2 * g.n1 + c;
  return (((size_t)(c >> g.lgW) * g.n2 + k2) << g.lgW) + (c & (g.W - 1));
}

// Forward transform of the own time columns held in `src` (layout lc*sT + r,
// own columns lc in [HW, HW+W)) into the buffer t1 (t1_index, twiddled).
// Returns the buffer (src or other) that holds the transforms: data read
// from the exchange may go only to the other one, since a thread can
// complete its reads while another is still reading these.
template <typename R>
OFX_DEV cpx<R>* fwd_cols_to_t1(const Geo<R>& g, cpx<R>* src, cpx<R>* other, u64* t1,
                               unsigned tag) {
  typedef cpx<R> C;
  typedef Tagged<C> TT;
  C* res = lds_fft<R, -1>(src + g.HW * g.sT, other + g.HW * g.sT, g.lg2, g.W, g.sT, g.tw2);
  const int tot = g.W * g.n2;
  for (int idx = OFX_TID; idx < tot; idx += OFX_NT) {
    const int cl = idx & (g.W - 1), k2 = idx >> g.lgW;
    const int c = g.c0 + cl;
    C v = res[cl * g.sT + k2];
    v = cmul(v, fourstep_tw(g, c * k2));
    const TagPtr t = tp_plane<C>(g.plain, t1, g.n, t1_index(g, c, k2));
    TT::put(const_cast<u64*>(t.p), t.ws, t.plain, v, tag);
  }
  return res - g.HW * g.sT;
}

// Own frequency rows from t1 into `dst` (layout k2l*sF + c), then length-n1
// forward transforms; returns the buffer holding bins (k2l*sF + k1).  `ok`
// (uniform over the workgroup) is false when the cluster gave up.
template <typename R>
OFX_DEV cpx<R>* t1_to_rows_fwd(const Geo<R>& g, const u64* t1, unsigned tag, cpx<R>* dst,
                               cpx<R>* other, Spin& sp, bool& ok) {
  typedef cpx<R> C;
  ok = read_tagged<C>(
      g.Rr * g.n1, tag,
      [&](int idx) {
        return tp_plane<C>(g.plain, t1, g.n, t1_index(g, idx & (g.n1 - 1), g.k20 + (idx >> g.lg1)));
      },
      [&](int idx, C v) { dst[(idx >> g.lg1) * g.sF + (idx & (g.n1 - 1))] = v; }, sp, g.plain);
  ok = wg_all(ok, g.plain);
  return lds_fft<R, -1>(dst, other, g.lg1, g.Rr, g.sF, g.tw1);
}

// After the exchange of an attempt: read the cluster's error records, this
// workgroup's t1 rows (into B1; A0 in B0 stays intact for a rejection) and,
// for RK23 (kx != null), the neighbours' K4 boundary columns into K1's halo,
// all polled together, then reduce the records (max).  *emax = max error
// quotient (+inf when any member saw a non-finite field).  Returns false
// (uniform over the workgroup) when the cluster gave up.
template <typename R>
OFX_DEV bool gather_err_and_rows(const ClusterDesc& cd, int par, const Geo<R>& g, const u64* t1,
                                 const u64* kx, cpx<R>* K1, unsigned tag, double* emax,
                                 double* red, Spin& sp, cpx<R>* rows, int rec_per_member) {
  typedef cpx<R> C;
  const int nw = slot_waves();
  const int nrec = rec_per_member;  // records per member: one per wave, or one
  const int ns = cd.C * nrec;
  const int nr = g.Rr * g.n1;
  const int nk = kx ? 6 * g.n2 : 0;
  const int n2 = g.n2;
  const int left = (g.rank + g.Cn - 1) & (g.Cn - 1), right = (g.rank + 1) & (g.Cn - 1);
  // kx element of halo entry e = side*3n2 + col*n2 + r (row shift across the wrap)
  auto kx_src = [&](int e) {
    const int sc3 = e >> g.lg2, r = e & (n2 - 1);
    const int side = sc3 / 3, col = sc3 - 3 * side;
    int rr = r, q;
    if (side == 0) {
      q = left * 2 + 1;
      if (g.c0 - 3 + col < 0) rr = (r - 1) & (n2 - 1);
    } else {
      q = right * 2 + 0;
      if (g.c0 + g.W + col >= g.n1) rr = (r + 1) & (n2 - 1);
    }
    return tp_plane<C>(g.plain, kx, (size_t)g.Cn * 6 * n2, (size_t)(q * 3 + col) * n2 + rr);
  };
  double v = -1.0;
  bool ok = read_tagged<C>(
      ns + nr + nk, tag,
      [&](int idx) -> TagPtr {
        if (idx < ns) return TagPtr{slot_at<R>(cd, par, idx & (g.Cn - 1), idx >> g.lgC, 0), 1, 0};
        idx -= ns;
        if (idx < nr)
          return tp_plane<C>(g.plain, t1, g.n, t1_index(g, idx & (g.n1 - 1), g.k20 + (idx >> g.lg1)));
        return kx_src(idx - nr);
      },
      [&](int idx, C x) {
        if (idx < ns) {
          v = red_op(v, unpack_d<R>(x), 0);
          return;
        }
        idx -= ns;
        if (idx < nr) {
          rows[(idx >> g.lg1) * g.sF + (idx & (g.n1 - 1))] = x;
          return;
        }
        idx -= nr;
        const int sc3 = idx >> g.lg2, r = idx & (n2 - 1);
        const int side = sc3 / 3, col = sc3 - 3 * side;
        const int lc = side == 0 ? 1 + col : g.HW + g.W + col;
        K1[lc * g.sT + r] = x;
      },
      sp, g.plain);
#ifdef OFX_EMU
  (void)red;
  (void)nw;
  *emax = v;
  return ok;
#else
  for (int o = 32; o > 0; o >>= 1) v = red_op(v, __shfl_xor(v, o), 0);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  ok = wg_all(ok, g.plain);
  double r = red[0];
  for (int w = 1; w < nw; w++) r = red_op(r, red[w], 0);
  *emax = r;
  return ok;
#endif
}

// Entry of sample (column c, row k2) in the transpose buffer t2.  Tagged
// exchanges (clusters of 64 and more, a row or two per member): row-major
// k2*n1 + c, so a member writes its rows as contiguous runs and the column
// owners read runs of LC samples per row.  Plain exchanges (small clusters,
// many rows per member): column-major c*n2 + k2, where both sides move runs
// of Rr and n2 samples.
template <typename R>
OFX_DEV size_t t2_index(const Geo<R>& g, int c, int k2) {
  return g.plain ? (size_t)c * g.n2 + k2 : (size_t)k2 * g.n1 + c;
}

// Inverse length-n1 transforms of the own rows, conjugate twiddle, write the
// buffer t2.  Returns the buffer (src or other) holding the transforms (see
// fwd_cols_to_t1).
template <typename R>
OFX_DEV cpx<R>* rows_inv_to_t2(const Geo<R>& g, cpx<R>* src, cpx<R>* other, u64* t2,
                               unsigned tag) {
  typedef cpx<R> C;
  typedef Tagged<C> TT;
  C* res = lds_fft<R, +1>(src, other, g.lg1, g.Rr, g.sF, g.tw1);
  const int tot = g.Rr * g.n1;
  for (int idx = OFX_TID; idx < tot; idx += OFX_NT) {
    const int k2l = idx & (g.Rr - 1), c = idx >> g.lgR;
    const int k2 = g.k20 + k2l;
    C v = res[k2l * g.sF + c];
    v = cmulc(v, fourstep_tw(g, c * k2));
    const TagPtr t = tp_plane<C>(g.plain, t2, g.n, t2_index(g, c, k2));
    TT::put(const_cast<u64*>(t.p), t.ws, t.plain, v, tag);
  }
  return res;
}

// Own+halo columns from t2, inverse length-n2 transforms, scale 1/N and
// place into A0 (B0) with the row shift that wrapped halo columns carry.
// Returns false (uniform over the workgroup) when the cluster gave up.
// dst: a buffer no thread still reads (not the one the inverse rows left
// their transforms in), other: the second work buffer.
template <typename R>
OFX_DEV bool t2_to_cols_inv(const Geo<R>& g, const u64* t2, unsigned tag, Spin& sp, cpx<R>* dst,
                            cpx<R>* other) {
  typedef cpx<R> C;
  const int tot = g.LC * g.n2;
  // entry idx: consecutive threads read consecutive t2 entries, i.e.
  // consecutive columns of one row (row-major t2: row k2 = idx / LC, local
  // column lc = idx % LC) or consecutive rows of one column (column-major)
  auto split = [&](int idx, int& lc, int& k2) {
    if (g.plain) {
      lc = idx >> g.lg2;
      k2 = idx & (g.n2 - 1);
    } else {
      k2 = idx / g.LC;
      lc = idx - k2 * g.LC;
    }
  };
  bool ok = read_tagged<C>(
      tot, tag,
      [&](int idx) {
        int lc, k2;
        split(idx, lc, k2);
        int gc = g.c0 - g.HW + lc;
        gc = gc < 0 ? gc + g.n1 : (gc >= g.n1 ? gc - g.n1 : gc);
        return tp_plane<C>(g.plain, t2, g.n, t2_index(g, gc, k2));
      },
      [&](int idx, C v) {
        int lc, k2;
        split(idx, lc, k2);
        dst[lc * g.sT + k2] = v;
      },
      sp, g.plain);
  ok = wg_all(ok, g.plain);
  C* res = lds_fft<R, +1>(dst, other, g.lg2, g.LC, g.sT, g.tw2);
  const R inv = R(1) / R(g.n);
  for (int idx = OFX_TID; idx < tot; idx += OFX_NT) {
    const int lc = idx >> g.lg2, rp = idx & (g.n2 - 1);
    const int gc = g.c0 - g.HW + lc;
    const int sh = gc < 0 ? 1 : (gc >= g.n1 ? -1 : 0);
    const int r = (rp + sh) & (g.n2 - 1);
    g.B0[lc * g.sT + r] = cscale(res[lc * g.sT + rp], inv);
  }
  OFX_SYNC();
  return ok;
}


// ------------------------------------------------------------------------
// One adaptive RK23 pass: the Bogacki-Shampine 3(2) step and the step
// control of ode23/OPO_CUDA_ODE23.cu:375-557 (restated in the oracle's
// orc_rk23), with FSAL (K1 <- K4 of the accepted step, computed before the
// absorber and the dispersion).  Buffers (halo HW = 4 columns):
//   B0 A0 (all columns), B1 E, B2 K2, B3 K3, B4 K1/K4.
// Validity per attempt: K1 [1, LC-1), K2 [2, LC-2), K3 and E [3, LC-3), K4
// on the own columns.  K4's three boundary columns on each side go through
// cd.kx at the first barrier and fill K1's halo for the next step.
// Returns ST_RUNNING or a terminal status; counts attempts / accepted steps.
// ------------------------------------------------------------------------
template <typename R, int MP, int XM>
OFX_DEV int rk23_pass(const LaunchArgs& a, const ClusterDesc& cd, const InstDesc& d, Geo<R>& g,
                      int rank, double* red, SyncState& sy, const double* breg,
                      const R* areg, long long& att, long long& acc) {
  typedef cpx<R> C;
  typedef Tagged<C> TT;
  const int n = g.n, n1 = g.n1, n2 = g.n2;
  const size_t kxs = (size_t)g.Cn * 6 * n2 * TT::NW;  // one parity half of kx
  const R cc = (R)(0.5 / d.dt / d.w0);
  C* A0 = g.B0;
  C* Eb = g.B1;
  C* K2 = g.B2;
  C* K3 = g.B3;
  C* K1 = g.B4;
  const int HW = g.HW, LC = g.LC, W = g.W;
  const int Ns = d.nsteps;
  const int sel = (int)d.sel;
  const int nchunks = Ns / sel;
  const double A1c = 1.0 / 2.0, A2c = 3.0 / 4.0;
  const R a1 = (R)A1c, a2 = (R)A2c, a3 = (R)(4.0 / 9.0), a4 = (R)(1.0 / 3.0),
          a5 = (R)(2.0 / 9.0);
  const R b1 = (R)(1.0 / 8.0), b2 = (R)(1.0 / 3.0), b3 = (R)(1.0 / 4.0), b4 = (R)(7.0 / 24.0);
  // K of the nonlinear term of (A0 + cin*kin) on columns [lo, hi)
  auto kstage = [&](C* out, const C* kin, R cin, int lo, int hi, double bdwz, R gk) {
    const C eb = expi(-bdwz, R(0));
    for (int idx = OFX_TID; idx < (hi - lo) * n2; idx += OFX_NT) {
      const int lc = lo + (idx >> g.lg2), r = idx & (n2 - 1);
      const int p = lc * g.sT + r;
      C sc = A0[p], su = A0[p + g.sT], sd = A0[p - g.sT];
      if (kin) {
        sc = cadd(sc, cscale(kin[p], cin));
        su = cadd(su, cscale(kin[p + g.sT], cin));
        sd = cadd(sd, cscale(kin[p - g.sT], cin));
      }
      const C nc = nl_term_e<R>(sc, cmul(wt_at(g, lc, r), eb));
      const C nu = nl_term_e<R>(su, cmul(wt_at(g, lc + 1, r), eb));
      const C nd = nl_term_e<R>(sd, cmul(wt_at(g, lc - 1, r), eb));
      out[p] = k_term<R>(nc, nu, nd, gk, cc);
    }
    OFX_SYNC();
  };
  double k = 0.0, stepprev = 4.0, stepmod = 4.0;
  kstage(K1, nullptr, R(0), 1, LC - 1, d.bdiffw0 * k * d.h, (R)(stepmod * d.h * gl(d.G)[0]));
  for (int chunk = 0; chunk < nchunks; chunk++) {
    stepmod = stepprev;
    const double cend = (double)(chunk + 1) * (double)sel;
    while (k < cend) {
      if ((int)k >= Ns) return ST_BAD_ARG;
      const double Gk = gl(d.G)[(int)k];
      const R gk = (R)(stepmod * d.h * Gk);
      // K2 from A1 = A0 + K1/2
      kstage(K2, K1, a1, 2, LC - 2, d.bdiffw0 * (k + A1c * stepmod) * d.h, gk);
      // K3 from A0 + 3/4 K2, then E = A0 + (4/9 K3 + 1/3 K2) + 2/9 K1
      {
        const C eb = expi(-(d.bdiffw0 * (k + A2c * stepmod) * d.h), R(0));
        for (int idx = OFX_TID; idx < (LC - 6) * n2; idx += OFX_NT) {
          const int lc = 3 + (idx >> g.lg2), r = idx & (n2 - 1);
          const int p = lc * g.sT + r;
          const C sc = cadd(A0[p], cscale(K2[p], a2));
          const C su = cadd(A0[p + g.sT], cscale(K2[p + g.sT], a2));
          const C sd = cadd(A0[p - g.sT], cscale(K2[p - g.sT], a2));
          const C nc = nl_term_e<R>(sc, cmul(wt_at(g, lc, r), eb));
          const C nu = nl_term_e<R>(su, cmul(wt_at(g, lc + 1, r), eb));
          const C nd = nl_term_e<R>(sd, cmul(wt_at(g, lc - 1, r), eb));
          const C k3 = k_term<R>(nc, nu, nd, gk, cc);
          K3[p] = k3;
          Eb[p] = cadd(cadd(A0[p], cadd(cscale(k3, a3), cscale(K2[p], a4))), cscale(K1[p], a5));
        }
        OFX_SYNC();
      }
      // own columns: K4 from NL(E), the 2nd-order solution, the error quotient;
      // K4 replaces K1 (FSAL), the absorbed E goes to K2 for the transforms
      double lmax = -1.0;
      int bad = 0;
      {
        const C eb = expi(-(d.bdiffw0 * (k + stepmod) * d.h), R(0));
        for (int idx = OFX_TID; idx < W * n2; idx += OFX_NT) {
          const int lc = HW + (idx >> g.lg2), r = idx & (n2 - 1);
          const int p = lc * g.sT + r;
          const C e = Eb[p];
          const C nc = nl_term_e<R>(e, cmul(wt_at(g, lc, r), eb));
          const C nu = nl_term_e<R>(Eb[p + g.sT], cmul(wt_at(g, lc + 1, r), eb));
          const C nd = nl_term_e<R>(Eb[p - g.sT], cmul(wt_at(g, lc - 1, r), eb));
          const C k4 = k_term<R>(nc, nu, nd, gk, cc);
          const C k1 = K1[p], k2 = K2[p], k3 = K3[p];
          const C a1s = cadd(cadd(cadd(A0[p], cadd(cscale(k4, b1), cscale(k3, b2))), cscale(k2, b3)),
                             cscale(k1, b4));
          const double ex = (double)e.x, ey = (double)e.y;
          const double aE = sqrt(ex * ex + ey * ey);
          if (aE > 0.0) {
            const double dx = (double)a1s.x - ex, dy = (double)a1s.y - ey;
            const double q = 0.5 * sqrt(dx * dx + dy * dy) / (aE + 1.0);
            if (q > lmax) lmax = q;
          }
          if (isnan(ex) || isnan(ey)) bad = 1;
          K1[p] = k4;
          const R f = g.ab[(lc - HW) * n2 + r];
          K2[p] = f != R(1) ? cscale(e, f) : e;
        }
        OFX_SYNC();
      }
      // exchange: error records, K4's boundary columns, forward column FFTs
      const unsigned x1 = ++sy.nx, tag1 = sy.base + x1;
      const int par = (int)(x1 & 1u);
      u64* t1 = (u64*)cd.t1 + (size_t)par * n * TT::NW;
      u64* kx = (u64*)cd.kx + par * kxs;
      publish_slots<R>(cd, par, rank, 0, bad ? INFINITY : fmin(lmax, DBL_MAX), 0.0, 0.0, tag1);
      for (int idx = OFX_TID; idx < 6 * n2; idx += OFX_NT) {
        const int sc3 = idx >> g.lg2, r = idx & (n2 - 1);
        const int side = sc3 / 3, col = sc3 - 3 * side;
        const int lc = side == 0 ? HW + col : HW + W - 3 + col;
        const TagPtr t = tp_plane<C>(g.plain, kx, (size_t)g.Cn * 6 * n2,
                                     ((size_t)(rank * 2 + side) * 3 + col) * n2 + r);
        TT::put(const_cast<u64*>(t.p), t.ws, t.plain, K1[lc * g.sT + r], tag1);
      }
      fwd_cols_to_t1<R>(g, K2, K3, t1, tag1);
      double emax;
      {
        Spin sp = spin_begin(cd, a.timeout_ticks);
        if (!xchg_ready<XM>(cd, sy, a.timeout_ticks) ||
            !gather_err_and_rows<R>(cd, par, g, t1, kx, K1, tag1, &emax, red, sp, g.B1,
                                    slot_waves()))
          return ST_BARRIER_TIMEOUT;
      }
      att++;
      if (isinf(emax)) return ST_NONFINITE;
      const double pmax = emax < 0.0 ? -1.0 : 4.0 * 100.0 * emax;
      const bool flag0 = pmax >= d.max_err;
      const bool flag1 = !flag0 && pmax >= d.min_err;
      if (flag0 && stepmod > 1.0 && (k - stepmod) > 0.0) {
        // reject: back by the failed step, one minimal step smaller, K1 from A0
        k = k - stepmod;
        stepmod = stepmod - 1.0;
        stepprev = stepmod;
        kstage(K1, nullptr, R(0), 1, LC - 1, d.bdiffw0 * (k + stepmod) * d.h,
               (R)(stepmod * d.h * gl(d.G)[(int)k]));
        k = k + stepmod;
        continue;
      }
      // ---------------- accept ----------------
      acc++;
      if (rank == 0 && d.stepmods && (int)k < Ns) gl(d.stepmods)[(int)k] = (float)stepmod;
      // (K1's halo already holds the neighbours' K4 boundary columns, read
      // with the rows)
      // frequency phase: rows, dispersion (step just taken), snapshot, inverse rows
      const double s_used = stepmod;
      C* F = lds_fft<R, -1>(g.B1, g.B2, g.lg1, g.Rr, g.sF, g.tw1);
      C* Fo = (F == g.B1) ? g.B2 : g.B1;
      auto disperse = [&](int idx, double beta, R al) {
        const int k2l = idx >> g.lg1, k1i = idx & (n1 - 1);
        const int kk = (g.k20 + k2l) + n2 * k1i;
        C v = F[k2l * g.sF + k1i];
        const C ph = expi(-beta * s_used, R(0));
        const R am = al == R(1) ? R(1) : (al == R(0) ? R(0) : (R)pow(al, (R)s_used));
        v = cscale(cmul(v, ph), am);
        F[k2l * g.sF + k1i] = v;
        if (d.apft && chunk < d.n_cols) stg(gl((C*)d.apft), (size_t)chunk * n + kk, v);
      };
#pragma unroll
      for (int q = 0; q < MP; q++) {
        const int idx = OFX_TID + q * OFX_NT;
        if (idx < g.Rr * n1) disperse(idx, breg[q], areg[q]);
      }
      for (int idx = OFX_TID + MP * OFX_NT; idx < g.Rr * n1; idx += OFX_NT) {
        const int kk = (g.k20 + (idx >> g.lg1)) + n2 * (idx & (n1 - 1));
        disperse(idx, gl(d.beta)[kk], (R)gl(d.alpha)[kk]);
      }
      OFX_SYNC();
      {
        const unsigned x2 = ++sy.nx, tag2 = sy.base + x2;
        u64* t2 = (u64*)cd.t2 + (size_t)(x2 & 1u) * n * TT::NW;
        C* ri = rows_inv_to_t2<R>(g, F, Fo, t2, tag2);
        Spin sp = spin_begin(cd, a.timeout_ticks);
        if (!xchg_ready<XM>(cd, sy, a.timeout_ticks) ||
            !t2_to_cols_inv<R>(g, t2, tag2, sp, ri == F ? Fo : F, ri))
          return ST_BARRIER_TIMEOUT;
      }
      // step control after an accepted step (OPO_CUDA_ODE23.cu order)
      if (k + stepmod >= cend && k < cend - 1.0) {
        const double stepend = (cend - k) - 1.0;
        stepprev = stepmod;
        stepmod = stepend;
      } else {
        if (!flag1 && !flag0 && (k + stepmod + stepmod) < cend - 1.0) stepmod = stepmod * 2.0;
        stepprev = stepmod;
      }
      k = k + stepmod;
    }
  }
  return ST_RUNNING;
}

// OFX_RKL (defined by the translation unit that builds the kernel's RK23
// variant, ofx_backend_rkl.hip): the RK23 pass reads the cluster's
// descriptor from a copy in LDS instead of global memory, where each use is
// a new load since the kernel's stores may alias it.  The copy is faster for
// the RK23 pass and slower for the Heun pass (whose register allocation it
// moves), so the variant lives in a code object of its own.
#ifdef OFX_RKL
#define OFX_RKL_PARAM , const ClusterDesc& cdr
#define OFX_RKL_ARG , cdr
#define OFX_RK23_CD cdr
#define OFX_CDL_PARAM , ClusterDesc* cdl
#else
#define OFX_RKL_PARAM
#define OFX_RKL_ARG
#define OFX_RK23_CD cd
#define OFX_CDL_PARAM
#endif

// ------------------------------------------------------------------------
// One instance: single pass (trips == 0) or round trips.
// ------------------------------------------------------------------------
template <typename R, int XM>
OFX_DEV void run_instance(const LaunchArgs& a, const ClusterDesc& cd, const InstDesc& d,
                          int rank, cpx<R>* lds, double* red, int* sflag,
                          SyncState& sy OFX_RKL_PARAM) {
  typedef cpx<R> C;
  typedef Tagged<C> TT;
  // converged in an earlier launch (and not rewritten since): nothing to do
  if (a.conv_tol > 0.0 && a.trips > 0 && d.status[0] == ST_CONVERGED) return;
  Geo<R> g;
  g.n = d.n;
  g.n1 = d.n1;
  g.n2 = d.n2;
  g.lg1 = d.lg1;
  g.lg2 = d.lg2;
  g.lgn = d.lg1 + d.lg2;
  g.Cn = cd.C;
  g.lgC = 0;
  while ((1 << g.lgC) < g.Cn) g.lgC++;
  g.rank = rank;
  g.W = g.n1 >> g.lgC;
  g.lgW = g.lg1 - g.lgC;
  g.Rr = g.n2 >> g.lgC;
  g.lgR = g.lg2 - g.lgC;
  g.HW = cd.halo;
  g.LC = g.W + 2 * g.HW;
  g.sT = g.n2 + 1;
  g.sF = g.n1 + 1;
  g.c0 = rank * g.W;
  g.k20 = rank * g.Rr;
  g.plain = XM == XM_BARRIER;
  {
    int b1 = g.LC * g.sT, b2 = g.Rr * g.sF;
    int bs = b1 > b2 ? b1 : b2;
    g.B0 = lds;
    g.B1 = lds + bs;
    g.B2 = lds + 2 * bs;
    g.B3 = lds + 3 * bs;  // RK23 only (the launch sizes LDS for 5 buffers then)
    g.B4 = lds + 4 * bs;
    // tables after the field buffers (the launch sizes LDS for them): the
    // transforms' pass twiddles of both lengths (n1 and n2 entries, the last
    // Ns0 unused) and the two factors of
    // exp(i w0 t_j) (j = h n1 + l: wt[j] = wt[h n1] wt[l], since t_j is
    // linear in j within each half of the fftshifted axis and n/2 is a
    // multiple of n1), 2 n1 + 2 n2 complex; then, 16-byte aligned and in
    // FP64, the two factors of the four-step twiddle (exp(-2 pi i q / n2),
    // q < n2, and exp(-2 pi i r / n), r < n1; their product is rounded once,
    // as a table entry would be); then the own columns' absorber factors
    // ((lc - HW) * n2 + r, W n2 reals)
    C* twl = lds + (d.method == METHOD_RK23 ? 5 : 3) * bs;
    const C* twg = (const C*)a.twiddle;
    const C* wtg = (const C*)d.wt;
    const int s1 = a.nmax >> d.lg1, s2 = a.nmax >> d.lg2;
    const int n1 = d.n1, n2 = d.n2;
    for (int j = OFX_TID; j < 2 * n1 + 2 * n2; j += OFX_NT) {
      C v = mkc<R>(R(0), R(0));
      if (j < n1) {
        if (j < n1 - pass_ns0(d.lg1)) v = twg[pass_twiddle_index(d.lg1, j) * s1];
      } else if (j < n1 + n2) {
        if (j - n1 < n2 - pass_ns0(d.lg2)) v = twg[pass_twiddle_index(d.lg2, j - n1) * s2];
      } else if (j < 2 * n1 + n2)
        v = wtg[j - n1 - n2];
      else
        v = wtg[(j - 2 * n1 - n2) * n1];
      twl[j] = v;
    }
    g.tw1 = twl;
    g.tw2 = twl + n1;
    g.wl = twl + n1 + n2;
    g.wh = twl + 2 * n1 + n2;
    // (the LDS base is 16-byte aligned; pointer arithmetic from it keeps the
    // accesses in the LDS address space)
    unsigned char* lbase = reinterpret_cast<unsigned char*>(lds);
    const size_t toff = (size_t)(reinterpret_cast<unsigned char*>(twl + 2 * n1 + 2 * n2) - lbase);
    cpx<double>* twd = reinterpret_cast<cpx<double>*>(lbase + ((toff + 15) & ~(size_t)15));
    {
      const double TWO_PI = 6.283185307179586232;
      for (int j = OFX_TID; j < n2 + n1; j += OFX_NT) {
        const double ang = j < n2 ? -TWO_PI * (double)j / (double)n2
                                  : -TWO_PI * (double)(j - n2) / (double)(n1 * n2);
        twd[j] = mkc<double>(cos(ang), sin(ang));
      }
    }
    g.twq = twd;
    g.twr = twd + n2;
    R* abl = reinterpret_cast<R*>(twd + n2 + n1);
    const R* abg = (const R*)d.absorb;
    for (int i = OFX_TID; i < g.W * n2; i += OFX_NT)
      abl[i] = abg[gidx(g, g.HW + (i >> g.lg2), i & (n2 - 1))];
    g.ab = abl;
  }
  const int n = g.n, n1 = g.n1, n2 = g.n2;
  C* field = (C*)d.field;
  // parity halves of the tagged transpose buffers
  auto t1_of = [&](unsigned x) { return (u64*)cd.t1 + (size_t)(x & 1u) * n * TT::NW; };
  auto t2_of = [&](unsigned x) { return (u64*)cd.t2 + (size_t)(x & 1u) * n * TT::NW; };
  const double dt = d.dt, w0 = d.w0;
  const R cc = (R)(0.5 / dt / w0);
  const int nsteps = d.nsteps;
  const double sel = d.sel;
  const double nChunks = (double)nsteps / sel;
  const bool cavity = a.trips > 0 && d.cav != nullptr;
  const int trips = a.trips > 0 ? a.trips : 1;
  int status = ST_RUNNING;
  long long attempts = 0, accepted = 0;
  const int trip_base = d.trips_done[0];
  const bool fresh = d.fresh[0] != 0;
  // Descriptor fields the step loop reads, copied once: the descriptor is in
  // global memory and the loop's stores may alias it, so every use inside
  // the loop would be a new load on the step's critical path (for the
  // coupling G(k), two dependent loads at the start of each step).
  const double* const Gd = uniform(d.G);
  const double bdiffw0 = uniform(d.bdiffw0), hmin = uniform(d.h);
  const double max_err = uniform(d.max_err), min_err = uniform(d.min_err);
  float* const stepmods = uniform(d.stepmods);
  C* const apft = uniform((C*)d.apft);
  const int n_cols = uniform(d.n_cols);

  // load the crystal-input field with halo columns
  for (int idx = OFX_TID; idx < g.LC * n2; idx += OFX_NT) {
    const int lc = idx >> g.lg2, r = idx & (n2 - 1);
    g.B0[lc * g.sT + r] = field[gidx(g, lc, r)];
  }
  OFX_SYNC();
  // Dispersion tables (beta, alpha) of this workgroup's own frequency bins,
  // held in registers for the whole instance (the thread->bin mapping of the
  // frequency phase is fixed); bins beyond MP per thread are read from memory.
  constexpr int MP = sizeof(R) == 4 ? 8 : 4;
  double breg[MP];
  R areg[MP];
#pragma unroll
  for (int q = 0; q < MP; q++) {
    const int idx = OFX_TID + q * OFX_NT;
    breg[q] = 0.0;
    areg[q] = R(0);
    if (idx < g.Rr * n1) {
      const int kk = (g.k20 + (idx >> g.lg1)) + n2 * (idx & (n1 - 1));
      breg[q] = gl(d.beta)[kk];
      areg[q] = (R)gl(d.alpha)[kk];
    }
  }

  int trip = 0;
  for (; trip < trips && status == ST_RUNNING; trip++) {
    long long trip_att = 0, trip_acc = 0;
    double k = 1.0, stepmod = 4.0;
    int chunk = 1;
    // In round trips every trip's step multipliers start from a zero row, as
    // the rows of StepSizeModifiers do (OpticalSim.m:172,237).  Rank 0 owns
    // them; every one of its threads writes each accepted entry, and thread
    // j clears (and at the trip's end copies) entries j, j + NT, ..., so each
    // thread's own program order already orders its clear, writes and copy.
    if (a.trips > 0 && rank == 0 && d.stepmods)
      for (int j = OFX_TID; j < nsteps; j += OFX_NT) gl(d.stepmods)[j] = 0.0f;
    if (d.method == METHOD_RK23) {
      status = rk23_pass<R, MP, XM>(a, OFX_RK23_CD, d, g, rank, red, sy, breg, areg, trip_att,
                                    trip_acc);
      attempts += trip_att;
      accepted += trip_acc;
      chunk = (int)nChunks + 1;  // skip the Heun loop
    }
    // ---------------- adaptive Heun pass (NEE_CPU.m:39-94) ----------------
    // G(k) of the step: loaded when the previous step was accepted, so the
    // load is in flight during that step's frequency phase and exchange
    double Gk = (double)chunk <= nChunks ? uniform(gl(Gd)[(int)k - 1]) : 0.0;
    while ((double)chunk <= nChunks) {
      const double bdwz0 = bdiffw0 * (k - 1.0) * hmin;
      const double bdwz1 = bdiffw0 * (k - 1.0 + stepmod) * hmin;
      const R gk = (R)(hmin * stepmod * Gk);
      C* A0 = g.B0;
      C* X = g.B1;
      C* K = g.B2;
      int bad = 0;
      // exp(i*(w0 t_j - bdwz)) = wt[j] * exp(-i*bdwz); bdwz reduced in FP64
      const C eb0 = expi(-bdwz0, R(0));
      const C eb1 = expi(-bdwz1, R(0));
      // Stage 1: K1 on columns 1..LC-2.  Each point evaluates the nonlinear
      // term at itself and its two neighbours j+-1 (the central difference of
      // NEE_CPU.m:104) instead of staging the term through LDS: three cheap
      // evaluations instead of a workgroup barrier and an LDS round trip.
      for (int idx = OFX_TID; idx < (g.LC - 2) * n2; idx += OFX_NT) {
        const int lc = 1 + (idx >> g.lg2), r = idx & (n2 - 1);
        const int p = lc * g.sT + r;
        const C v = A0[p];
        if (lc >= g.HW && lc < g.HW + g.W && !(isfinite((double)v.x) && isfinite((double)v.y)))
          bad = 1;
        const C nc = nl_term_e<R>(v, cmul(wt_at(g, lc, r), eb0));
        const C nu = nl_term_e<R>(A0[p + g.sT], cmul(wt_at(g, lc + 1, r), eb0));
        const C nd = nl_term_e<R>(A0[p - g.sT], cmul(wt_at(g, lc - 1, r), eb0));
        K[p] = k_term<R>(nc, nu, nd, gk, cc);
      }
      OFX_SYNC();
      // Stage 2 on own columns: K2 from NL(A1 = A0 + K1) at j and j+-1, the
      // corrector E = A0 + K1/2 + K2/2, the error quotient and the absorber.
      // E goes to X (neighbours still read K1).
      double lmax = -1.0;
      auto corrector = [&](int idx, R f) {
        const int lc = g.HW + (idx >> g.lg2), r = idx & (n2 - 1);
        const int p = lc * g.sT + r;
        const C k1 = K[p];
        const C a0 = A0[p];
        const C a1 = cadd(a0, k1);
        const C nc = nl_term_e<R>(a1, cmul(wt_at(g, lc, r), eb1));
        const C nu = nl_term_e<R>(cadd(A0[p + g.sT], K[p + g.sT]), cmul(wt_at(g, lc + 1, r), eb1));
        const C nd = nl_term_e<R>(cadd(A0[p - g.sT], K[p - g.sT]), cmul(wt_at(g, lc - 1, r), eb1));
        const C k2 = k_term<R>(nc, nu, nd, gk, cc);
        C e = cadd(cadd(a0, cscale(k1, R(0.5))), cscale(k2, R(0.5)));
        const double ex = (double)e.x - (double)a1.x, ey = (double)e.y - (double)a1.y;
        const double err = sqrt(ex * ex + ey * ey) /
                           (sqrt((double)e.x * (double)e.x + (double)e.y * (double)e.y) + 1.0);
        if (err > lmax) lmax = err;
        if (f != R(1)) e = cscale(e, f);
        X[p] = e;
      };
      for (int idx = OFX_TID; idx < g.W * n2; idx += OFX_NT)
        corrector(idx, g.ab[idx]);
      // the workgroup's error maximum (+inf for a non-finite field): each
      // wave's before the barrier, then one record for the whole member, so
      // the cluster reads C records instead of C times the waves
      double wmax = bad ? INFINITY : fmin(lmax, DBL_MAX);
#ifndef OFX_EMU
      for (int o = 32; o > 0; o >>= 1) wmax = red_op(wmax, __shfl_xor(wmax, o), 0);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wmax;
#endif
      OFX_SYNC();
      // exchange: error records and the speculative forward column transforms of E
      const unsigned x1 = ++sy.nx, tag1 = sy.base + x1;
      const int par = (int)(x1 & 1u);
      if (OFX_TID == 0) {
#ifndef OFX_EMU
        for (int w = 1; w < slot_waves(); w++) wmax = red_op(wmax, red[w], 0);
#endif
        Tagged<C>::put(slot_at<R>(cd, par, rank, 0, 0), 1, 0, pack_d<R>(wmax), tag1);
      }
      C* cols = fwd_cols_to_t1<R>(g, X, K, t1_of(x1), tag1);
      C* rows = cols == X ? K : X;
      double gmax;
      {
        Spin sp = spin_begin(cd, a.timeout_ticks);
        if (!xchg_ready<XM>(cd, sy, a.timeout_ticks) ||
            !gather_err_and_rows<R>(cd, par, g, t1_of(x1), nullptr, nullptr, tag1, &gmax, red,
                                    sp, rows, 1)) {
          status = ST_BARRIER_TIMEOUT;
          break;
        }
      }
      const bool gbad = isinf(gmax);
      attempts++;
      trip_att++;
      if (gbad) {
        status = ST_NONFINITE;
        break;
      }
      const double pcterr = 100.0 * gmax;
      if (pcterr > max_err && stepmod > 1.0) {
        stepmod -= 1.0;
        continue;  // reject: A0 is untouched in B0
      }
      // ---------------- accept ----------------
      accepted++;
      trip_acc++;
      if (rank == 0 && stepmods) gl(stepmods)[(int)k - 1] = (float)stepmod;
      const double s_used = stepmod;
      double ns = stepmod;
      if (pcterr < min_err) ns += 1.0;
      if (k < nsteps && (k + ns) > nsteps) ns = (double)nsteps - k;
      const double knew = k + ns;
      // the next step's G(k) (index clamped: past the last step it is not used)
      const double Gnext = gl(Gd)[((int)knew < nsteps ? (int)knew : nsteps) - 1];
      int cend = chunk;  // chunks [chunk, cend) end with this step
      while ((double)cend <= nChunks && !(knew < cend * sel && knew <= nsteps)) cend++;
      // frequency phase: rows, dispersion, snapshots, inverse rows
      C* F = lds_fft<R, -1>(rows, cols, g.lg1, g.Rr, g.sF, g.tw1);
      C* Fo = (F == rows) ? cols : rows;
      auto disperse = [&](int idx, double beta, R al) {
        const int k2l = idx >> g.lg1, k1 = idx & (n1 - 1);
        const int kk = (g.k20 + k2l) + n2 * k1;
        C v = F[k2l * g.sF + k1];
        const C ph = expi(-beta * s_used, R(0));
        const R am = al == R(1) ? R(1) : (al == R(0) ? R(0) : (R)pow(al, (R)s_used));
        v = cscale(cmul(v, ph), am);
        F[k2l * g.sF + k1] = v;
        if (apft) {
          OFX_GLOBAL C* ap = gl(apft);
          for (int cc2 = chunk; cc2 < cend; cc2++)
            if (cc2 - 1 < n_cols) stg(ap, (size_t)(cc2 - 1) * n + kk, v);
        }
      };
#pragma unroll
      for (int q = 0; q < MP; q++) {
        const int idx = OFX_TID + q * OFX_NT;
        if (idx < g.Rr * n1) disperse(idx, breg[q], areg[q]);
      }
      for (int idx = OFX_TID + MP * OFX_NT; idx < g.Rr * n1; idx += OFX_NT) {
        const int kk = (g.k20 + (idx >> g.lg1)) + n2 * (idx & (n1 - 1));
        disperse(idx, gl(d.beta)[kk], (R)gl(d.alpha)[kk]);
      }
      OFX_SYNC();
      {
        const unsigned x2 = ++sy.nx, tag2 = sy.base + x2;
        C* ri = rows_inv_to_t2<R>(g, F, Fo, t2_of(x2), tag2);
        Spin sp = spin_begin(cd, a.timeout_ticks);
        if (!xchg_ready<XM>(cd, sy, a.timeout_ticks) ||
            !t2_to_cols_inv<R>(g, t2_of(x2), tag2, sp, ri == F ? Fo : F, ri)) {
          status = ST_BARRIER_TIMEOUT;
          break;
        }
      }
      k = knew;
      stepmod = ns;
      chunk = cend;
      Gk = uniform(Gnext);
    }
    if (status != ST_RUNNING) break;
    // A plain pass (trips == 0) is not a round trip: it leaves the trip
    // count, the trip log and the history ring to the round trips.
    if (a.trips > 0 && rank == 0 && d.sm_hist && d.stepmods) {
      float* row = d.sm_hist + (size_t)((trip_base + trip) % d.hist_cap) * (size_t)nsteps;
      for (int j = OFX_TID; j < nsteps; j += OFX_NT) gl(row)[j] = gl(d.stepmods)[j];
    }
    if (a.trips > 0 && rank == 0 && d.trip_log && trip_base + trip < d.trip_cap) {
      double* tl = d.trip_log + 4 * (size_t)(trip_base + trip);
      tl[2] = (double)trip_att;
      tl[3] = (double)trip_acc;
    }
    if (!cavity) continue;
    // ---------------- cavity: OC split, detuning, pump (one FFT pair) ----------------
    const unsigned xc1 = ++sy.nx, tagc1 = sy.base + xc1;
    C* ccols = fwd_cols_to_t1<R>(g, g.B0, g.B1, t1_of(xc1), tagc1);
    C* crows = ccols == g.B0 ? g.B1 : g.B0;
    {
      bool ok = xchg_ready<XM>(cd, sy, a.timeout_ticks);
      C* F = crows;
      if (ok) {
        Spin sp = spin_begin(cd, a.timeout_ticks);
        F = t1_to_rows_fwd<R>(g, t1_of(xc1), tagc1, crows, ccols, sp, ok);
      }
      if (!ok) {
        status = ST_BARRIER_TIMEOUT;
        break;
      }
      C* Fo = (F == crows) ? ccols : crows;
      const C* pump = (const C*)d.pump;
      const C* cav = (const C*)d.cav;
      const C* ocm = (const C*)d.ocm;
      C* ocs = (C*)d.oc_spec;
      double* esd = d.esd_sum;
      double* xesd = d.xesd_sum;
      C* hist = d.hist ? (C*)d.hist + (size_t)((trip_base + trip) % d.hist_cap) * (size_t)n : nullptr;
      const bool have_prev = (trip_base + trip) > 0 && !(fresh && trip == 0);
      double acc[3] = {0.0, 0.0, 0.0};  // sum (S-Sp)^2, sum S^2, sum S
      for (int idx = OFX_TID; idx < g.Rr * n1; idx += OFX_NT) {
        const int k2l = idx >> g.lg1, k1 = idx & (n1 - 1);
        const int kk = (g.k20 + k2l) + n2 * k1;
        C v = F[k2l * g.sF + k1];
        xesd[kk] += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        if (hist) hist[kk] = v;
        C o = cmul(v, ocm[kk]);
        const double S = (double)o.x * (double)o.x + (double)o.y * (double)o.y;
        if (have_prev) {
          C op = ocs[kk];
          const double Sp = (double)op.x * (double)op.x + (double)op.y * (double)op.y;
          acc[0] += (S - Sp) * (S - Sp);
        }
        acc[1] += S * S;
        acc[2] += S;
        ocs[kk] = o;
        esd[kk] += S;  // each kk belongs to one thread of one member
        C m = cmul(cav[kk], expi(-d.delay * dw_of(kk, n, dt), R(0)));
        F[k2l * g.sF + k1] = cadd(cmul(v, m), pump ? pump[kk] : mkc<R>(R(0), R(0)));
      }
      OFX_SYNC();
      const unsigned xc2 = ++sy.nx, tagc2 = sy.base + xc2;
      const int par = (int)(xc2 & 1u);
      publish_slots<R>(cd, par, rank, 1, acc[0], acc[1], acc[2], tagc2);
      C* cri = rows_inv_to_t2<R>(g, F, Fo, t2_of(xc2), tagc2);
      double tot[3];
      ok = xchg_ready<XM>(cd, sy, a.timeout_ticks);
      if (ok) {
        Spin sp = spin_begin(cd, a.timeout_ticks);
        // (the columns go to B0, so their transforms run in B1 and B2)
        C* cd1 = cri == g.B1 ? g.B2 : g.B1;
        ok = t2_to_cols_inv<R>(g, t2_of(xc2), tagc2, sp, cd1, cd1 == g.B1 ? g.B2 : g.B1);
      }
      if (!ok || !gather_sums<R>(cd, par, tagc2, tot, red, a.timeout_ticks)) {
        status = ST_BARRIER_TIMEOUT;
        break;
      }
      const double conv = have_prev ? (tot[1] > 0.0 ? sqrt(tot[0] / tot[1]) : 0.0) : INFINITY;
      if (rank == 0 && d.trip_log && trip_base + trip < d.trip_cap) {
        double* tl = d.trip_log + 4 * (size_t)(trip_base + trip);
        tl[0] = conv;
        tl[1] = tot[2];
      }
      if (a.conv_tol > 0.0 && conv < a.conv_tol) {
        status = ST_CONVERGED;
        trip++;
        break;
      }
    }
  }
  if (status == ST_RUNNING) status = ST_DONE;
  // store own columns (after a barrier so no member still reads the input halo)
  if (status != ST_BARRIER_TIMEOUT) {
    sy.nreal++;
    if (!cluster_barrier(cd, sy.nreal * (unsigned)g.Cn, a.timeout_ticks, sflag))
      status = ST_BARRIER_TIMEOUT;
  }
  if (status != ST_BARRIER_TIMEOUT) {
    for (int idx = OFX_TID; idx < g.W * n2; idx += OFX_NT) {
      const int lc = g.HW + (idx >> g.lg2), r = idx & (n2 - 1);
      field[gidx(g, lc, r)] = g.B0[lc * g.sT + r];
    }
  }
  if (rank == 0 && OFX_TID == 0) {
    d.status[0] = status;
    if (a.trips > 0) {
      if (trip > 0) d.fresh[0] = 0;
      d.trips_done[0] += trip;
    }
    d.counters[0] += attempts;
    d.counters[1] += accepted;
  }
}

// (with OFX_RKL, cdl: workgroup-shared storage for the descriptor's copy)
template <typename R, int XM>
OFX_DEV void cluster_main(const LaunchArgs& a, int wg, cpx<R>* lds, double* red,
                          int* sflag OFX_CDL_PARAM) {
  const int ci = a.wg_cluster[wg];
  if (ci < 0) return;
  const ClusterDesc& cd = a.clusters[ci];
#ifdef OFX_RKL
  if (OFX_TID == 0) *cdl = cd;
  OFX_SYNC();
  const ClusterDesc& cdr = *cdl;
#endif
  const int rank = a.wg_rank[wg];
  SyncState sy;
#ifdef OFX_EMU
  sy.base = __atomic_load_n(cd.epoch, __ATOMIC_RELAXED);
#else
  sy.base = __hip_atomic_load(cd.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
  sy.nx = 0;
  sy.nreal = 0;
  sy.sflag = sflag;
#ifndef OFX_EMU
  // placement diagnostic: hardware XCC id of every member
  if (threadIdx.x == 0) {
    const int x = (int)__builtin_amdgcn_s_getreg(OFX_HWREG_XCC_ID);
    __hip_atomic_store(&cd.xcc[rank], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a.xcc_ids) a.xcc_ids[wg] = x;
  }
#endif
  for (int ii = 0; ii < cd.n_inst; ii++) {
    const int id = a.inst_order[cd.first_inst + ii];
    run_instance<R, XM>(a, cd, a.inst[id], rank, lds, red, sflag, sy OFX_RKL_ARG);
    if (flag_load(cd.abort_flag)) break;
  }
  // the tags this member used end at base + nx: the next launch starts past
  // the largest over the cluster
  if (OFX_TID == 0) {
#ifdef OFX_EMU
    unsigned cur = __atomic_load_n(cd.epoch + 1, __ATOMIC_RELAXED);
    while (cur < sy.nx &&
           !__atomic_compare_exchange_n(cd.epoch + 1, &cur, sy.nx, false, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED)) {
    }
#else
    __hip_atomic_fetch_max(cd.epoch + 1, sy.nx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
  }
}

}  // namespace ofx
