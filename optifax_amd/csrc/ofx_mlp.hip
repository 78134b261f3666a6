// Spectral surrogate of one crystal pass: MLP training and inference on gfx950.
//
// Network (PyTorch Linear convention, weight [out][in]):
//   z1 = x W0^T + b0, h1 = gelu(z1); z2 = h1 W1^T + b1, h2 = gelu(z2);
//   z3 = h2 W2^T + b2, h3 = gelu(z3); y = h3 W3^T + b3
// Loss: mean over the batch of the relative L2 error ||y - t|| / ||t||.
//
// Precision: GEMM operands BF16 (weights, activations, output gradients),
// accumulation FP32 on the matrix cores (v_mfma_f32_32x32x16_bf16); master
// weights, pre-activations, weight gradients and Adam moments FP32.
//
// Every product is one "NT" GEMM, C[m][n] = sum_k A[m][k] B[n][k], with both
// operands contiguous in k, so tiles stage into LDS with 16-byte loads and
// stores and feed the MFMA without transposition:
//   forward        Z   = H  . Wt^T      A = H   [batch][in],   B = Wt  [out][in]
//   data grad      dH  = dZ . Wn^T      A = dZ  [batch][out],  B = Wn  [in][out]
//   weight grad    dWt = dZ^T . H       A = dZT [out][batch],  B = HT  [in][batch]
// which needs a BF16 copy of each weight in both orientations (Wt, Wn; built
// after each optimiser step) and transposed copies of H and dZ (written by
// the producing GEMM's epilogue, 8-byte stores of 4 consecutive rows).
//
// Tiles: 64 x 64 outputs per workgroup of 4 waves (each wave one 32 x 32 MFMA
// block), K in slices of 64 double-buffered in LDS, with the global loads of
// the slices two steps ahead in flight during the current slice's MFMAs.
// For the 512-row batch and width 2048, a layer is 8 x 32 = 256 workgroups:
// one per CU.
//
// Training step: the weight-gradient GEMM's epilogue applies Adam to the
// FP32 master weights and writes both BF16 copies, so the gradient itself
// never reaches memory; the bias gradients and their Adam step are one small
// kernel per layer.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/optifax_hip.h"

namespace ofxm {

typedef __bf16 bf16;
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum { EPI_GELU = 0, EPI_BIAS = 1, EPI_DGELU = 2, EPI_F32 = 3, EPI_ADAM = 4 };

constexpr int TM = 64, TN = 64, TK = 64;
constexpr int LDK = TK + 8;  // LDS row stride (bf16): 144 B, staggers the 16-B reads over banks

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

// A pointer that is the same in every lane, marked as such (so that buffer
// descriptors built from it live in scalar registers)
__device__ __forceinline__ void* uniform_ptr(const void* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (void*)(((uint64_t)hi << 32) | lo);
}

// Adam (Kingma & Ba) hyper-parameters of one step; c1, c2 = bias corrections
struct AdamHP {
  float lr, b1, b2, eps, c1, c2;
};

// one Adam update of parameter p with gradient g (moments m, v updated in place)
__device__ __forceinline__ float adam_one(float p, float g, float& m, float& v, const AdamHP& h) {
  const float mi = h.b1 * m + (1.f - h.b1) * g;
  const float vi = h.b2 * v + (1.f - h.b2) * g * g;
  m = mi;
  v = vi;
  return p - h.lr * (mi / h.c1) / (sqrtf(vi / h.c2) + h.eps);
}

struct GemmArgs {
  const bf16* A;  // [M][lda], k-contiguous
  const bf16* B;  // [N][ldb], k-contiguous
  int M, N, K, lda, ldb;
  const float* bias;  // [N] (GELU, BIAS)
  float* Cf;          // FP32 out [M][ldc]: Z (GELU), y (BIAS), dW (F32)
  int ldc;
  bf16* Cb;           // BF16 out [M][ldcb]: H (GELU), dZ (DGELU)
  int ldcb;
  bf16* Ct;           // BF16 transposed out [N][ldct]: H^T, dZ^T (optional)
  int ldct;
  const float* Zin;   // pre-activation [M][ldz] (DGELU)
  int ldz;
  // ADAM: the product is the gradient of the FP32 master weights Cf; their
  // moments am, av share Cf's layout; Cb and Ct receive the BF16 copies
  float* am;
  float* av;
  AdamHP hp;
};

// Epilogue of a 64 x 64 output tile (all 512 threads of gemm_nt).  The MFMA
// layout scatters a lane's outputs over rows, so the tile goes through LDS
// and the epilogue runs on rows: thread t owns row t / 8, columns
// 8 (t % 8) .. + 7, and moves FP32 data as 32-byte and BF16 data as 16-byte
// pieces (a row of the tile is 256 contiguous bytes of an FP32 matrix).  A
// transposed BF16 output goes back through LDS once more and is stored as
// 16-byte column pieces of 8 rows.
//   GELU:  Cf = Z = acc + bias, Cb = H = gelu(Z), Ct = H^T
//   BIAS:  Cf = acc + bias
//   DGELU: Cb = dZ = acc * gelu'(Zin), Ct = dZ^T
//   F32:   Cf = acc
//   ADAM:  acc is the gradient of the FP32 master weights Cf (moments am, av):
//          one Adam step, then Cb = BF16 copy, Ct = its transpose
//   xch: half 1's accumulators (written before the call, 16 KB);
//   T:   64 x 68 floats, then 64 x 72 BF16 (the B stages, free after the K loop)
template <int EPI>
__device__ __forceinline__ void tile_epilogue(const GemmArgs& g, f32x16 acc, const float* xch,
                                              float* T, int half, int th, int m0, int n0) {
  constexpr int LT = 68;
  const int lane = th & 63, w = th >> 6, wr = w >> 1, wc = w & 1;
  if (half == 0) {
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] += xch[i * 256 + th];
    // accumulator i of lane l holds row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32
    const int col = 32 * wc + (lane & 31), rb = 32 * wr + 4 * (lane >> 5);
#pragma unroll
    for (int i = 0; i < 16; i++) T[(rb + 8 * (i >> 2) + (i & 3)) * LT + col] = acc[i];
  }
  __syncthreads();
  const int t = th + 256 * half;
  const int r = t >> 3, c8 = (t & 7) * 8;
  const int row = m0 + r, col = n0 + c8;
  float v[8];
  {
    const float4* gp = reinterpret_cast<const float4*>(T + r * LT + c8);
    const float4 a = gp[0], b = gp[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  }
  auto ld8 = [](const float* q, float (&o)[8]) {
    const float4 a = reinterpret_cast<const float4*>(q)[0], b = reinterpret_cast<const float4*>(q)[1];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  };
  auto st8 = [](float* q, const float (&o)[8]) {
    reinterpret_cast<float4*>(q)[0] = make_float4(o[0], o[1], o[2], o[3]);
    reinterpret_cast<float4*>(q)[1] = make_float4(o[4], o[5], o[6], o[7]);
  };
  bf16x8 ob;
  const size_t ic = (size_t)row * g.ldc + col;
  if (EPI == EPI_GELU || EPI == EPI_BIAS) {
    float bv[8];
    ld8(g.bias + col, bv);
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] += bv[e];
    st8(g.Cf + ic, v);
    if (EPI == EPI_BIAS) return;
#pragma unroll
    for (int e = 0; e < 8; e++) ob[e] = (bf16)gelu(v[e]);
  } else if (EPI == EPI_DGELU) {
    float z[8];
    ld8(g.Zin + (size_t)row * g.ldz + col, z);
#pragma unroll
    for (int e = 0; e < 8; e++) ob[e] = (bf16)(v[e] * gelu_d(z[e]));
  } else if (EPI == EPI_ADAM) {
    float pv[8], mv[8], vv[8];
    ld8(g.Cf + ic, pv);
    ld8(g.am + ic, mv);
    ld8(g.av + ic, vv);
#pragma unroll
    for (int e = 0; e < 8; e++) {
      pv[e] = adam_one(pv[e], v[e], mv[e], vv[e], g.hp);
      ob[e] = (bf16)pv[e];
    }
    st8(g.Cf + ic, pv);
    st8(g.am + ic, mv);
    st8(g.av + ic, vv);
  } else {
    st8(g.Cf + ic, v);
    return;
  }
  *reinterpret_cast<bf16x8*>(g.Cb + (size_t)row * g.ldcb + col) = ob;
  if (!g.Ct) return;  // (uniform: no barrier is skipped by part of the workgroup)
  bf16* Wb = reinterpret_cast<bf16*>(T);  // 64 x 72 BF16 over the (consumed) FP32 tile
  __syncthreads();
  *reinterpret_cast<bf16x8*>(Wb + r * 72 + c8) = ob;
  __syncthreads();
  // transposed copy: thread t writes column t / 8 of the tile, rows 8 (t % 8) .. + 7
  const int cc = t >> 3, r8 = (t & 7) * 8;
  bf16x8 tv;
#pragma unroll
  for (int e = 0; e < 8; e++) tv[e] = Wb[(r8 + e) * 72 + cc];
  *reinterpret_cast<bf16x8*>(g.Ct + (size_t)(n0 + cc) * g.ldct + m0 + r8) = tv;
}

// 512 threads = two halves of 4 waves.  Half h accumulates the K slices
// kt = 2i + h of the same 64 x 64 tile (its own double-buffered LDS), so a CU
// holds two waves per SIMD with independent load streams; at the end half 1
// hands its accumulators to half 0 through LDS, which assembles the tile for
// the epilogue that both halves run.
struct GemmSmem {
  bf16 As[2][2][TM * LDK];  // [half][stage]
  bf16 Bs[2][2][TN * LDK];
};

template <int EPI>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int bx, int by, GemmSmem& sm) {
  auto& As = sm.As;
  auto& Bs = sm.Bs;
  // (the half index is the same across a wave; readfirstlane lets the compiler
  // keep what depends on it, such as the slice offsets, in scalar registers)
  const int tid = threadIdx.x, th = tid & 255;
  const int half = __builtin_amdgcn_readfirstlane(tid >> 8);
  const int lane = th & 63, w = th >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int m0 = by * TM, n0 = bx * TN;
  const int r = th >> 3, c = (th & 7) * 8;
  // Tile loads through buffer descriptors of the 64-row A and B panels:
  // loop-invariant per-thread byte offsets plus the slice offset in an
  // SGPR, so no address registers are rewritten while loads are in flight
  // (with 64-bit per-thread addresses the compiler reused registers of
  // in-flight loads and had to wait for them)
  const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
      uniform_ptr(g.A + (size_t)m0 * g.lda), (short)0,
      __builtin_amdgcn_readfirstlane(TM * g.lda * 2), 0x00020000);
  const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
      uniform_ptr(g.B + (size_t)n0 * g.ldb), (short)0,
      __builtin_amdgcn_readfirstlane(TN * g.ldb * 2), 0x00020000);
  const int oa0 = 2 * (r * g.lda + c), oa1 = 2 * ((r + 32) * g.lda + c);
  const int ob0 = 2 * (r * g.ldb + c), ob1 = 2 * ((r + 32) * g.ldb + c);
  const int so0 = r * LDK + c, so1 = (r + 32) * LDK + c;
  const int nk = g.K / TK;
  const int iters = (nk + 1) >> 1;  // both halves run the same number of iterations
  // Global loads run two of this half's slices ahead of the MFMAs: slice
  // kt + 2 waits in one register set while slice kt + 4 is in flight in the
  // other, so each load has two iterations to arrive.
  bf16x8 s0[4], s1[4];
  auto load = [&](int kt, bf16x8 (&r)[4]) {
    const int so = kt * TK * 2;
    r[0] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rA, oa0, so, 0));
    r[1] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rA, oa1, so, 0));
    r[2] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rB, ob0, so, 0));
    r[3] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rB, ob1, so, 0));
  };
  auto stash = [&](int stage, const bf16x8 (&r)[4]) {
    *reinterpret_cast<bf16x8*>(&As[half][stage][so0]) = r[0];
    *reinterpret_cast<bf16x8*>(&As[half][stage][so1]) = r[1];
    *reinterpret_cast<bf16x8*>(&Bs[half][stage][so0]) = r[2];
    *reinterpret_cast<bf16x8*>(&Bs[half][stage][so1]) = r[3];
  };
  // (slices past the end are clamped to the last one: the loads and LDS
  // stores stay unconditional, which keeps the compiler's wait counts exact;
  // their MFMAs are skipped)
  const int kl = nk - 1;
  load(min(half, kl), s0);
  stash(0, s0);
  load(min(half + 2, kl), s0);
  __syncthreads();
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; i++) acc[i] = 0.0f;
  // MFMA operands: lane l supplies row (A) / column (B) l % 32, k = 8 (l / 32) .. + 7
  const int ao = (32 * wr + (lane & 31)) * LDK + 8 * (lane >> 5);
  const int bo = (32 * wc + (lane & 31)) * LDK + 8 * (lane >> 5);
  // iteration it: MFMAs on slice kt = 2 it + half (LDS stage it & 1); `held`
  // holds slice kt + 2, `next` receives slice kt + 4
  auto step = [&](int it, bf16x8 (&held)[4], bf16x8 (&next)[4]) {
    const int cur = it & 1;
    const int kt = 2 * it + half;
    load(min(kt + 4, kl), next);
    if (kt < nk) {
      const bf16* as = &As[half][cur][ao];
      const bf16* bs = &Bs[half][cur][bo];
#pragma unroll
      for (int ks = 0; ks < TK; ks += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(as + ks);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(bs + ks);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
    stash(cur ^ 1, held);
    __syncthreads();
  };
  for (int it = 0; it < iters; it += 2) {
    step(it, s0, s1);
    if (it + 1 < iters) step(it + 1, s1, s0);
  }
  // half 1 -> half 0 through LDS (the A buffers: 256 threads x 16 floats = 16 KB)
  float* xch = reinterpret_cast<float*>(&As[0][0][0]);
  if (half == 1) {
#pragma unroll
    for (int i = 0; i < 16; i++) xch[i * 256 + th] = acc[i];
  }
  __syncthreads();
  tile_epilogue<EPI>(g, acc, xch, reinterpret_cast<float*>(&Bs[0][0][0]), half, th, m0, n0);
}

template <int EPI>
__global__ void __launch_bounds__(512) gemm_nt(GemmArgs g) {
  __shared__ __align__(16) GemmSmem sm;
  gemm_tile<EPI>(g, blockIdx.x, blockIdx.y, sm);
}

// Two independent products in one launch (1-D grid): workgroups below n1 take
// the tiles of g1, the others those of g2.  Pairs a data-gradient product
// (bound by operand traffic) with a weight-gradient product whose Adam
// epilogue is bound by HBM, so the CUs work on both at once.
template <int E1, int E2>
__global__ void __launch_bounds__(512) gemm_nt2(GemmArgs g1, GemmArgs g2, int n1) {
  __shared__ __align__(16) GemmSmem sm;
  const int b = blockIdx.x;
  if (b < n1) {
    const int nx = g1.N / TN;
    gemm_tile<E1>(g1, b % nx, b / nx, sm);
  } else {
    const int nx = g2.N / TN;
    gemm_tile<E2>(g2, (b - n1) % nx, (b - n1) / nx, sm);
  }
}

// Relative-L2 loss of each row; dL/dy = (y - t) / (B ||y - t|| ||t||) as BF16
// (row-major; the transpose is a separate tiled pass); n a multiple of 64
// (16-byte aligned rows, 4 values per thread and load).  Target row r is
// T[idx[r]] (a device dataset) or T[r] when idx is null.  One workgroup
// (256 threads) per row.
__global__ void __launch_bounds__(256) rel_l2_loss(const float* __restrict__ y,
                                                   const float* __restrict__ T,
                                                   const int* __restrict__ idx, int n, float inv_b,
                                                   bf16* __restrict__ dy,
                                                   float* __restrict__ row_loss) {
  __shared__ float red[2][4];
  const int r = blockIdx.x;
  const float* yr = y + (size_t)r * n;
  const float* tr = T + (size_t)(idx ? idx[r] : r) * n;
  float se = 0.f, st = 0.f;
  for (int i = 4 * threadIdx.x; i < n; i += 1024) {
    const float4 a = *reinterpret_cast<const float4*>(yr + i);
    const float4 b = *reinterpret_cast<const float4*>(tr + i);
    const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    se += (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
    st += (b.x * b.x + b.y * b.y) + (b.z * b.z + b.w * b.w);
  }
  for (int o = 32; o > 0; o >>= 1) {
    se += __shfl_xor(se, o);
    st += __shfl_xor(st, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = se;
    red[1][threadIdx.x >> 6] = st;
  }
  __syncthreads();
  const float E = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  const float Tn = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  const float ne = sqrtf(E), nt = Tn > 0.f ? sqrtf(Tn) : 1.f;
  if (threadIdx.x == 0) row_loss[r] = ne / nt;
  const float sc = ne > 0.f ? inv_b / (ne * nt) : 0.f;
  for (int i = 4 * threadIdx.x; i < n; i += 1024) {
    const float4 a = *reinterpret_cast<const float4*>(yr + i);
    const float4 b = *reinterpret_cast<const float4*>(tr + i);
    bf16x4 o;
    o[0] = (bf16)(sc * (a.x - b.x));
    o[1] = (bf16)(sc * (a.y - b.y));
    o[2] = (bf16)(sc * (a.z - b.z));
    o[3] = (bf16)(sc * (a.w - b.w));
    *reinterpret_cast<bf16x4*>(dy + (size_t)r * n + i) = o;
  }
}

// dst [cols][rows] = src [rows][cols]^T (BF16), 64 x 64 tiles through LDS;
// rows and cols multiples of 64.  Thread t moves 16 consecutive values of
// one row in and 16 of one column out, as 16-byte pieces.
__global__ void __launch_bounds__(256) transpose_bf16(const bf16* __restrict__ src, int rows,
                                                      int cols, bf16* __restrict__ dst) {
  __shared__ bf16 tile[64][72];   // 144-byte rows: 16-byte aligned, staggered banks
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int a = threadIdx.x >> 2, b = (threadIdx.x & 3) * 16;
  const bf16* sp = src + (size_t)(r0 + a) * cols + c0 + b;
  *reinterpret_cast<bf16x8*>(&tile[a][b]) = *reinterpret_cast<const bf16x8*>(sp);
  *reinterpret_cast<bf16x8*>(&tile[a][b + 8]) = *reinterpret_cast<const bf16x8*>(sp + 8);
  __syncthreads();
  bf16x8 o0, o1;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    o0[e] = tile[b + e][a];
    o1[e] = tile[b + 8 + e][a];
  }
  bf16* dp = dst + (size_t)(c0 + a) * rows + r0 + b;
  *reinterpret_cast<bf16x8*>(dp) = o0;
  *reinterpret_cast<bf16x8*>(dp + 8) = o1;
}

// Bias gradients of all four layers, db[n] = sum_b dZT[n][b] (one wave per n,
// 8 BF16 of the row per lane and load; cols a multiple of 8), and the mean of
// the row losses (the last workgroup, one wave, fixed order) into hist[slot].
// adam != 0: the biases take their Adam step with moments am, av; else db
// goes to out.
struct BiasJobs {
  const bf16* dT[4];
  float* out[4];
  float* p[4];
  float* am[4];
  float* av[4];
  int rows[4];
  int blk0[5];  // first workgroup of each layer; blk0[4] = the loss-mean workgroup
  int cols, adam, slot;
  AdamHP hp;
  const float* row_loss;
  float* hist;
};

__global__ void __launch_bounds__(256) bias_and_loss(BiasJobs j) {
  const int b = blockIdx.x, lane = threadIdx.x & 63;
  if (b == j.blk0[4]) {
    if (threadIdx.x >= 64) return;
    float s = 0.f;
    for (int i = lane; i < j.cols; i += 64) s += j.row_loss[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) j.hist[j.slot] = s / (float)j.cols;
    return;
  }
  int l = 0;
  while (l < 3 && b >= j.blk0[l + 1]) l++;
  const int n = (b - j.blk0[l]) * 4 + (threadIdx.x >> 6);
  if (n >= j.rows[l]) return;
  const bf16* dT = j.dT[l];
  float s = 0.f;
  for (int c = 8 * lane; c < j.cols; c += 512) {
    const bf16x8 v = *reinterpret_cast<const bf16x8*>(dT + (size_t)n * j.cols + c);
#pragma unroll
    for (int e = 0; e < 8; e++) s += (float)v[e];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane) return;
  if (j.adam) {
    float mm = j.am[l][n], vv = j.av[l][n];
    j.p[l][n] = adam_one(j.p[l][n], s, mm, vv, j.hp);
    j.am[l][n] = mm;
    j.av[l][n] = vv;
  } else {
    j.out[l][n] = s;
  }
}

// BF16 copies of one weight matrix W [rows][cols] (FP32): Wt = W, Wn = W^T
// (optional).  64 x 64 tiles through LDS; rows, cols multiples of 64.
__global__ void __launch_bounds__(256) weight_copies(const float* __restrict__ W, int rows,
                                                     int cols, bf16* __restrict__ Wt,
                                                     bf16* __restrict__ Wn) {
  __shared__ bf16 tile[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  for (int i = threadIdx.x; i < 4096; i += 256) {
    const int rr = i >> 6, cc = i & 63;
    const bf16 v = (bf16)W[(size_t)(r0 + rr) * cols + c0 + cc];
    Wt[(size_t)(r0 + rr) * cols + c0 + cc] = v;
    tile[rr][cc] = v;
  }
  if (!Wn) return;
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += 256) {
    const int cc = i >> 6, rr = i & 63;
    Wn[(size_t)(c0 + cc) * rows + r0 + rr] = tile[rr][cc];
  }
}

// input rows (FP32, n_in wide) -> BF16 [batch][ld_in] and [ld_in][batch], zero
// padded; rows come from a device dataset through idx (or in order when idx
// is null).  A workgroup takes 64 rows x LI_COLS columns: thread t reads 16
// consecutive values of row t / 4 in each 64-column half (8-byte loads when
// n_in is even, so every row starts on an 8-byte boundary), writes them as
// 16-byte BF16 pieces and, through LDS, 16-byte pieces of the transpose.
constexpr int LI_COLS = 128;
__global__ void __launch_bounds__(256) load_inputs(const float* __restrict__ X,
                                                   const int* __restrict__ idx, int batch, int n_in,
                                                   int ld_in, bf16* __restrict__ h0,
                                                   bf16* __restrict__ h0T) {
  __shared__ bf16 tile[64][LI_COLS + 8];
  const int r0 = blockIdx.y * 64, cb = blockIdx.x * LI_COLS;
  const int rr = threadIdx.x >> 2, q = (threadIdx.x & 3) * 16;
  const float* xr = X + (idx ? (size_t)idx[r0 + rr] : (size_t)(r0 + rr)) * n_in;
  const bool pairs = (n_in & 1) == 0;
#pragma unroll
  for (int hh = 0; hh < LI_COLS; hh += 64) {
    const int cl = hh + q, c = cb + cl;
    if (c >= ld_in) break;
    float v[16];
#pragma unroll
    for (int e = 0; e < 16; e += 2) {
      if (pairs && c + e + 1 < n_in) {
        const float2 p = *reinterpret_cast<const float2*>(xr + c + e);
        v[e] = p.x;
        v[e + 1] = p.y;
      } else {
        v[e] = c + e < n_in ? xr[c + e] : 0.f;
        v[e + 1] = c + e + 1 < n_in ? xr[c + e + 1] : 0.f;
      }
    }
    bf16x8 o0, o1;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      o0[e] = (bf16)v[e];
      o1[e] = (bf16)v[8 + e];
    }
    bf16* hp = h0 + (size_t)(r0 + rr) * ld_in + c;
    *reinterpret_cast<bf16x8*>(hp) = o0;
    *reinterpret_cast<bf16x8*>(hp + 8) = o1;
    *reinterpret_cast<bf16x8*>(&tile[rr][cl]) = o0;
    *reinterpret_cast<bf16x8*>(&tile[rr][cl + 8]) = o1;
  }
  __syncthreads();
  // column cl of the block, rows rh .. rh + 31: four 16-byte pieces of h0T
  const int cl = threadIdx.x >> 1, rh = (threadIdx.x & 1) * 32;
  if (cb + cl >= ld_in) return;
  bf16* tp = h0T + (size_t)(cb + cl) * batch + r0 + rh;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = tile[rh + 8 * k + e][cl];
    *reinterpret_cast<bf16x8*>(tp + 8 * k) = o;
  }
}

}  // namespace ofxm

using namespace ofxm;

struct ofx_mlp {
  int device = 0;
  int n_in = 0, ld_in = 0, width = 0, n_out = 0, max_batch = 0;
  int kin[4] = {0, 0, 0, 0}, kout[4] = {0, 0, 0, 0};  // kin[0] is ld_in on the device
  size_t np_dev = 0, np_pub = 0;                      // parameters: device (padded) / public
  size_t off_w[4] = {0, 0, 0, 0}, off_b[4] = {0, 0, 0, 0};
  float *p = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;
  bf16* wt[4] = {nullptr, nullptr, nullptr, nullptr};  // [out][in]
  bf16* wn[4] = {nullptr, nullptr, nullptr, nullptr};  // [in][out], layers 1..3
  bf16* h[4] = {nullptr, nullptr, nullptr, nullptr};   // h[0] = input
  bf16* hT[4] = {nullptr, nullptr, nullptr, nullptr};
  float* z[4] = {nullptr, nullptr, nullptr, nullptr};  // z[1..3]
  bf16* dz[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // dz[4] = dL/dy
  bf16* dzT[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float *y = nullptr, *t = nullptr, *row_loss = nullptr, *hist = nullptr;
  float *xin = nullptr;  // host-batch staging (FP32 inputs)
  int hist_cap = 0;
  float *dsX = nullptr, *dsT = nullptr;
  int ds_n = 0;
  int* idx = nullptr;
  int idx_cap = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0.0;
  std::string err;
};

static int mlp_fail(ofx_mlp* m, const char* what, hipError_t e) {
  if (m) m->err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? OFX_E_NOMEM : OFX_E_HIP;
}
#define MCHK(m, call)                                        \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return mlp_fail((m), #call, _e);   \
  } while (0)
#define MLAUNCH(m) MCHK(m, hipGetLastError())

static int arg_err(ofx_mlp* m, const char* msg) {
  if (m) m->err = msg;
  return OFX_E_ARG;
}

template <int EPI>
static int gemm(ofx_mlp* m, const GemmArgs& a) {
  // every dimension is a multiple of the tile; checked once at create
  if (a.M % TM || a.N % TN || a.K % TK || a.M <= 0 || a.N <= 0 || a.K <= 0)
    return arg_err(m, "GEMM dimensions are not multiples of 64");
  hipLaunchKernelGGL(gemm_nt<EPI>, dim3(a.N / TN, a.M / TM), dim3(512), 0, m->stream, a);
  MLAUNCH(m);
  return OFX_OK;
}

template <int E1, int E2>
static int gemm2(ofx_mlp* m, const GemmArgs& a, const GemmArgs& b) {
  for (const GemmArgs* x : {&a, &b})
    if (x->M % TM || x->N % TN || x->K % TK || x->M <= 0 || x->N <= 0 || x->K <= 0)
      return arg_err(m, "GEMM dimensions are not multiples of 64");
  const int n1 = (a.N / TN) * (a.M / TM), n2 = (b.N / TN) * (b.M / TM);
  hipLaunchKernelGGL((gemm_nt2<E1, E2>), dim3(n1 + n2), dim3(512), 0, m->stream, a, b, n1);
  MLAUNCH(m);
  return OFX_OK;
}

static int refresh_weights(ofx_mlp* m) {
  for (int l = 0; l < 4; l++) {
    hipLaunchKernelGGL(weight_copies, dim3(m->kin[l] / 64, m->kout[l] / 64), dim3(256), 0,
                       m->stream, m->p + m->off_w[l], m->kout[l], m->kin[l], m->wt[l], m->wn[l]);
    MLAUNCH(m);
  }
  return OFX_OK;
}

static int forward(ofx_mlp* m, int B) {
  for (int l = 0; l < 4; l++) {
    GemmArgs a{};
    a.A = m->h[l];
    a.lda = m->kin[l];
    a.B = m->wt[l];
    a.ldb = m->kin[l];
    a.M = B;
    a.N = m->kout[l];
    a.K = m->kin[l];
    a.bias = m->p + m->off_b[l];
    int st;
    if (l < 3) {
      a.Cf = m->z[l + 1];
      a.ldc = m->width;
      a.Cb = m->h[l + 1];
      a.ldcb = m->width;
      a.Ct = m->hT[l + 1];
      a.ldct = B;
      st = gemm<EPI_GELU>(m, a);
    } else {
      a.Cf = m->y;
      a.ldc = m->n_out;
      st = gemm<EPI_BIAS>(m, a);
    }
    if (st) return st;
  }
  return OFX_OK;
}

// Loss (mean into hist[slot]) and gradients.  hp null: the gradients go to
// m->g (device layout).  hp set: each layer's weights and biases take their
// Adam step as soon as their gradients exist (in the weight-gradient GEMM's
// epilogue, which also writes the BF16 copies), so m->g is not touched; the
// layer's data gradient is computed first, from the weights before the step.
// Targets: row r is T[idx[r]] (idx null: T[r]).
static int backward(ofx_mlp* m, int B, int slot, const AdamHP* hp, const float* T,
                    const int* idx) {
  hipLaunchKernelGGL(rel_l2_loss, dim3(B), dim3(256), 0, m->stream, m->y, T, idx, m->n_out,
                     1.0f / (float)B, m->dz[4], m->row_loss);
  MLAUNCH(m);
  hipLaunchKernelGGL(transpose_bf16, dim3(m->n_out / 64, B / 64), dim3(256), 0, m->stream,
                     m->dz[4], B, m->n_out, m->dzT[4]);
  MLAUNCH(m);
  GemmArgs dh[4] = {}, dw[4] = {};
  for (int l = 3; l >= 0; l--) {
    if (l > 0) {
      // dZ_l = (dZ_{l+1} . Wn_l^T) * gelu'(z_l) : A = dz[l+1] [B][out], B = wn[l] [in][out]
      GemmArgs& d = dh[l];
      d.A = m->dz[l + 1];
      d.lda = m->kout[l];
      d.B = m->wn[l];
      d.ldb = m->kout[l];
      d.M = B;
      d.N = m->kin[l];
      d.K = m->kout[l];
      d.Zin = m->z[l];
      d.ldz = m->width;
      d.Cb = m->dz[l];
      d.ldcb = m->width;
      d.Ct = m->dzT[l];
      d.ldct = B;
    }
    // dWt_l [out][in] = dZ^T . H : A = dzT[l+1] [out][B], B = hT[l] [in][B]
    GemmArgs& a = dw[l];
    a.A = m->dzT[l + 1];
    a.lda = B;
    a.B = m->hT[l];
    a.ldb = B;
    a.M = m->kout[l];
    a.N = m->kin[l];
    a.K = B;
    a.ldc = m->kin[l];
    if (hp) {
      a.Cf = m->p + m->off_w[l];
      a.am = m->am + m->off_w[l];
      a.av = m->av + m->off_w[l];
      a.hp = *hp;
      a.Cb = m->wt[l];
      a.ldcb = m->kin[l];
      a.Ct = m->wn[l];  // null for layer 0
      a.ldct = m->kout[l];
    } else {
      a.Cf = m->g + m->off_w[l];
    }
  }
  // Launch order: dZ_3 alone; then each layer's weight gradient (and Adam
  // step) paired with the next data gradient, which reads other weights; the
  // last launch pairs layers 1 and 0.  dZ_l is always computed before layer
  // l's weights change.
  int st = gemm<EPI_DGELU>(m, dh[3]);
  if (st) return st;
  if (hp) {
    st = gemm2<EPI_DGELU, EPI_ADAM>(m, dh[2], dw[3]);
    if (!st) st = gemm2<EPI_DGELU, EPI_ADAM>(m, dh[1], dw[2]);
    if (!st) st = gemm2<EPI_ADAM, EPI_ADAM>(m, dw[1], dw[0]);
  } else {
    st = gemm2<EPI_DGELU, EPI_F32>(m, dh[2], dw[3]);
    if (!st) st = gemm2<EPI_DGELU, EPI_F32>(m, dh[1], dw[2]);
    if (!st) st = gemm2<EPI_F32, EPI_F32>(m, dw[1], dw[0]);
  }
  if (st) return st;
  // biases of every layer (their gradients do not feed any other) and the mean loss
  BiasJobs j{};
  int blk = 0;
  for (int l = 0; l < 4; l++) {
    const size_t ob = m->off_b[l];
    j.dT[l] = m->dzT[l + 1];
    j.out[l] = m->g + ob;
    j.p[l] = m->p + ob;
    j.am[l] = m->am + ob;
    j.av[l] = m->av + ob;
    j.rows[l] = m->kout[l];
    j.blk0[l] = blk;
    blk += (m->kout[l] + 3) / 4;
  }
  j.blk0[4] = blk;
  j.cols = B;
  j.adam = hp ? 1 : 0;
  if (hp) j.hp = *hp;
  j.slot = slot;
  j.row_loss = m->row_loss;
  j.hist = m->hist;
  hipLaunchKernelGGL(bias_and_loss, dim3(blk + 1), dim3(256), 0, m->stream, j);
  MLAUNCH(m);
  return OFX_OK;
}

// hyper-parameters of the next Adam step (advances the step count)
static AdamHP next_adam(ofx_mlp* m, float lr) {
  m->adam_t++;
  const double b1 = 0.9, b2 = 0.999;
  AdamHP h;
  h.lr = lr;
  h.b1 = (float)b1;
  h.b2 = (float)b2;
  h.eps = 1e-8f;
  h.c1 = (float)(1.0 - pow(b1, (double)m->adam_t));
  h.c2 = (float)(1.0 - pow(b2, (double)m->adam_t));
  return h;
}

// host batch -> device inputs (and targets)
static int stage_batch(ofx_mlp* m, int B, const float* x, const float* t) {
  MCHK(m, hipMemcpyAsync(m->xin, x, sizeof(float) * (size_t)B * m->n_in, hipMemcpyHostToDevice,
                         m->stream));
  hipLaunchKernelGGL(load_inputs, dim3((m->ld_in + LI_COLS - 1) / LI_COLS, B / 64), dim3(256), 0,
                     m->stream, m->xin,
                     (const int*)nullptr, B,
                     m->n_in, m->ld_in, m->h[0], m->hT[0]);
  MLAUNCH(m);
  if (t)
    MCHK(m, hipMemcpyAsync(m->t, t, sizeof(float) * (size_t)B * m->n_out, hipMemcpyHostToDevice,
                           m->stream));
  return OFX_OK;
}

static int check_batch(ofx_mlp* m, int B) {
  if (B < 64 || B > m->max_batch || B % 64) return arg_err(m, "batch must be a multiple of 64 in [64, max_batch]");
  return OFX_OK;
}

// public <-> device parameter layouts differ only in layer 0's padded columns
static int copy_params(ofx_mlp* m, float* dst_dev, const float* src_host, float* dst_host,
                       const float* src_dev) {
  const size_t w0 = (size_t)m->width * m->n_in;
  const hipMemcpyKind k = src_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  if (src_host) {
    MCHK(m, hipMemsetAsync(dst_dev, 0, sizeof(float) * (size_t)m->width * m->ld_in, m->stream));
    MCHK(m, hipMemcpy2DAsync(dst_dev, sizeof(float) * m->ld_in, src_host, sizeof(float) * m->n_in,
                             sizeof(float) * m->n_in, m->width, k, m->stream));
    MCHK(m, hipMemcpyAsync(dst_dev + m->off_b[0], src_host + w0,
                           sizeof(float) * (m->np_pub - w0), k, m->stream));
  } else {
    MCHK(m, hipMemcpy2DAsync(dst_host, sizeof(float) * m->n_in, src_dev, sizeof(float) * m->ld_in,
                             sizeof(float) * m->n_in, m->width, k, m->stream));
    MCHK(m, hipMemcpyAsync(dst_host + w0, src_dev + m->off_b[0],
                           sizeof(float) * (m->np_pub - w0), k, m->stream));
  }
  MCHK(m, hipStreamSynchronize(m->stream));
  return OFX_OK;
}

extern "C" {

void ofx_mlp_destroy(ofx_mlp* m);

int ofx_mlp_create(int device, int n_in, int width, int n_out, int max_batch, ofx_mlp** out) {
  if (!out) return OFX_E_ARG;
  *out = nullptr;
  if (n_in < 1 || width < 64 || width % 64 || n_out < 64 || n_out % 64 || max_batch < 64 ||
      max_batch % 64)
    return OFX_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OFX_E_NODEVICE;
  if (device < 0 || device >= ndev) return OFX_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return OFX_E_HIP;
  ofx_mlp* m = new ofx_mlp();
  m->device = device;
  m->n_in = n_in;
  m->ld_in = (n_in + 63) / 64 * 64;
  m->width = width;
  m->n_out = n_out;
  m->max_batch = max_batch;
  const int kin[4] = {m->ld_in, width, width, width};
  const int kout[4] = {width, width, width, n_out};
  size_t off = 0, pub = 0;
  for (int l = 0; l < 4; l++) {
    m->kin[l] = kin[l];
    m->kout[l] = kout[l];
    m->off_w[l] = off;
    off += (size_t)kin[l] * kout[l];
    m->off_b[l] = off;
    off += kout[l];
    pub += (size_t)(l == 0 ? n_in : kin[l]) * kout[l] + kout[l];
  }
  m->np_dev = off;
  m->np_pub = pub;
  hipError_t e = hipSuccess;
  auto al = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, bytes);
  };
  const size_t Bm = (size_t)max_batch;
  al((void**)&m->p, 4 * off);
  al((void**)&m->g, 4 * off);
  al((void**)&m->am, 4 * off);
  al((void**)&m->av, 4 * off);
  for (int l = 0; l < 4; l++) {
    al((void**)&m->wt[l], 2 * (size_t)kin[l] * kout[l]);
    if (l > 0) al((void**)&m->wn[l], 2 * (size_t)kin[l] * kout[l]);
    al((void**)&m->h[l], 2 * Bm * kin[l]);
    al((void**)&m->hT[l], 2 * Bm * kin[l]);
  }
  for (int l = 1; l < 4; l++) {
    al((void**)&m->z[l], 4 * Bm * width);
    al((void**)&m->dz[l], 2 * Bm * width);
    al((void**)&m->dzT[l], 2 * Bm * width);
  }
  al((void**)&m->dz[4], 2 * Bm * n_out);
  al((void**)&m->dzT[4], 2 * Bm * n_out);
  al((void**)&m->y, 4 * Bm * n_out);
  al((void**)&m->t, 4 * Bm * n_out);
  al((void**)&m->row_loss, 4 * Bm);
  al((void**)&m->xin, 4 * Bm * n_in);
  m->hist_cap = 1024;
  al((void**)&m->hist, 4 * (size_t)m->hist_cap);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&m->e0);
  if (e == hipSuccess) e = hipEventCreate(&m->e1);
  if (e == hipSuccess) e = hipMemsetAsync(m->p, 0, 4 * off, m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(m->am, 0, 4 * off, m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(m->av, 0, 4 * off, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    const int code = e == hipErrorOutOfMemory ? OFX_E_NOMEM : OFX_E_HIP;
    ofx_mlp_destroy(m);
    return code;
  }
  if (refresh_weights(m) || hipStreamSynchronize(m->stream) != hipSuccess) {
    ofx_mlp_destroy(m);
    return OFX_E_HIP;
  }
  *out = m;
  return OFX_OK;
}

void ofx_mlp_destroy(ofx_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* ptrs[] = {m->p,      m->g,      m->am,     m->av,     m->wt[0],  m->wt[1],  m->wt[2],
                  m->wt[3],  m->wn[1],  m->wn[2],  m->wn[3],  m->h[0],   m->h[1],   m->h[2],
                  m->h[3],   m->hT[0],  m->hT[1],  m->hT[2],  m->hT[3],  m->z[1],   m->z[2],
                  m->z[3],   m->dz[1],  m->dz[2],  m->dz[3],  m->dz[4],  m->dzT[1], m->dzT[2],
                  m->dzT[3], m->dzT[4], m->y,      m->t,      m->row_loss, m->xin,  m->hist,
                  m->dsX,    m->dsT,    m->idx};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (m->e0) (void)hipEventDestroy(m->e0);
  if (m->e1) (void)hipEventDestroy(m->e1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* ofx_mlp_last_error(const ofx_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

long long ofx_mlp_num_params(const ofx_mlp* m) { return m ? (long long)m->np_pub : 0; }

int ofx_mlp_set_params(ofx_mlp* m, const float* params) {
  if (!m || !params) return OFX_E_ARG;
  MCHK(m, hipSetDevice(m->device));
  int st = copy_params(m, m->p, params, nullptr, nullptr);
  if (st) return st;
  MCHK(m, hipMemsetAsync(m->am, 0, 4 * m->np_dev, m->stream));
  MCHK(m, hipMemsetAsync(m->av, 0, 4 * m->np_dev, m->stream));
  m->adam_t = 0;
  st = refresh_weights(m);
  if (st) return st;
  MCHK(m, hipStreamSynchronize(m->stream));
  return OFX_OK;
}

int ofx_mlp_get_params(ofx_mlp* m, float* params) {
  if (!m || !params) return OFX_E_ARG;
  MCHK(m, hipSetDevice(m->device));
  return copy_params(m, nullptr, nullptr, params, m->p);
}

int ofx_mlp_forward(ofx_mlp* m, int batch, const float* x, float* y) {
  if (!m || !x || !y) return OFX_E_ARG;
  int st = check_batch(m, batch);
  if (st) return st;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipEventRecord(m->e0, m->stream));
  st = stage_batch(m, batch, x, nullptr);
  if (!st) st = forward(m, batch);
  if (st) return st;
  MCHK(m, hipMemcpyAsync(y, m->y, sizeof(float) * (size_t)batch * m->n_out, hipMemcpyDeviceToHost,
                         m->stream));
  MCHK(m, hipEventRecord(m->e1, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  float ms = 0.f;
  MCHK(m, hipEventElapsedTime(&ms, m->e0, m->e1));
  m->last_ms = ms;
  return OFX_OK;
}

int ofx_mlp_gradients(ofx_mlp* m, int batch, const float* x, const float* t, float* loss,
                      float* grads) {
  if (!m || !x || !t) return OFX_E_ARG;
  int st = check_batch(m, batch);
  if (st) return st;
  MCHK(m, hipSetDevice(m->device));
  st = stage_batch(m, batch, x, t);
  if (!st) st = forward(m, batch);
  if (!st) st = backward(m, batch, 0, nullptr, m->t, nullptr);
  if (st) return st;
  if (loss) MCHK(m, hipMemcpyAsync(loss, m->hist, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (grads) return copy_params(m, nullptr, nullptr, grads, m->g);
  return OFX_OK;
}

int ofx_mlp_step(ofx_mlp* m, int batch, const float* x, const float* t, float lr, float* loss) {
  if (!m || !x || !t || !(lr >= 0.f)) return OFX_E_ARG;
  int st = check_batch(m, batch);
  if (st) return st;
  MCHK(m, hipSetDevice(m->device));
  st = stage_batch(m, batch, x, t);
  if (!st) st = forward(m, batch);
  if (!st) {
    const AdamHP hp = next_adam(m, lr);
    st = backward(m, batch, 0, &hp, m->t, nullptr);
  }
  if (st) return st;
  if (loss) MCHK(m, hipMemcpyAsync(loss, m->hist, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return OFX_OK;
}

int ofx_mlp_upload_dataset(ofx_mlp* m, int n, const float* X, const float* T) {
  if (!m || n < 1 || !X || !T) return OFX_E_ARG;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (m->dsX) (void)hipFree(m->dsX);
  if (m->dsT) (void)hipFree(m->dsT);
  m->dsX = m->dsT = nullptr;
  m->ds_n = 0;
  MCHK(m, hipMalloc((void**)&m->dsX, sizeof(float) * (size_t)n * m->n_in));
  MCHK(m, hipMalloc((void**)&m->dsT, sizeof(float) * (size_t)n * m->n_out));
  MCHK(m, hipMemcpy(m->dsX, X, sizeof(float) * (size_t)n * m->n_in, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->dsT, T, sizeof(float) * (size_t)n * m->n_out, hipMemcpyHostToDevice));
  m->ds_n = n;
  return OFX_OK;
}

// `steps` Adam steps on rows idx[s*batch .. s*batch+batch-1] of the uploaded
// dataset (every index in [0, n)); losses[s] = mean batch loss at step s
// (before its update).  The device time of the span is kept for ofx_mlp_last_ms.
int ofx_mlp_train(ofx_mlp* m, int batch, int steps, const int* idx, float lr, float* losses) {
  if (!m || !idx || steps < 1 || !(lr >= 0.f)) return OFX_E_ARG;
  int st = check_batch(m, batch);
  if (st) return st;
  if (!m->dsX) return arg_err(m, "no dataset uploaded");
  for (size_t i = 0; i < (size_t)steps * batch; i++)
    if (idx[i] < 0 || idx[i] >= m->ds_n) return arg_err(m, "dataset index out of range");
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (m->idx_cap < steps * batch) {
    if (m->idx) (void)hipFree(m->idx);
    m->idx = nullptr;
    m->idx_cap = 0;
    MCHK(m, hipMalloc((void**)&m->idx, sizeof(int) * (size_t)steps * batch));
    m->idx_cap = steps * batch;
  }
  if (m->hist_cap < steps) {
    (void)hipFree(m->hist);
    m->hist = nullptr;
    m->hist_cap = 0;
    MCHK(m, hipMalloc((void**)&m->hist, sizeof(float) * (size_t)steps));
    m->hist_cap = steps;
  }
  MCHK(m, hipMemcpyAsync(m->idx, idx, sizeof(int) * (size_t)steps * batch, hipMemcpyHostToDevice,
                         m->stream));
  MCHK(m, hipEventRecord(m->e0, m->stream));
  for (int s = 0; s < steps; s++) {
    const int* ix = m->idx + (size_t)s * batch;
    hipLaunchKernelGGL(load_inputs, dim3((m->ld_in + LI_COLS - 1) / LI_COLS, batch / 64),
                       dim3(256), 0, m->stream,
                       m->dsX, ix, batch, m->n_in, m->ld_in, m->h[0], m->hT[0]);
    MLAUNCH(m);
    st = forward(m, batch);
    if (!st) {
      const AdamHP hp = next_adam(m, lr);
      st = backward(m, batch, s, &hp, m->dsT, ix);
    }
    if (st) return st;
  }
  MCHK(m, hipEventRecord(m->e1, m->stream));
  MCHK(m, hipEventSynchronize(m->e1));
  float ms = 0.f;
  MCHK(m, hipEventElapsedTime(&ms, m->e0, m->e1));
  m->last_ms = ms;
  if (losses)
    MCHK(m, hipMemcpy(losses, m->hist, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost));
  return OFX_OK;
}

double ofx_mlp_last_ms(const ofx_mlp* m) { return m ? m->last_ms : 0.0; }

}  // extern "C"
