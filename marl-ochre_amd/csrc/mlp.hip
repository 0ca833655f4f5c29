// mlp.hip — the learned kinetics surrogate's network on gfx950 MFMA.
//
// A 4-layer MLP (80 -> 256 -> 256 -> 256 -> 28, SiLU on hidden layers) with
// FP32 master weights, computed with BF16 MFMA (v_mfma_f32_32x32x16_bf16) and
// FP32 accumulation.  Training uses one tiled GEMM kernel for the forward,
// the input-gradient and the weight-gradient products (operand transposes as
// template flags) with fused epilogues, plus bias-gradient, SiLU-derivative,
// loss and Adam kernels.  Inference uses a fused kernel (all four layers, the
// activations of a 64-cell tile kept in LDS, weights pre-packed per MFMA lane).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "marl_internal.h"
#include "mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// f32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N], f32 in memory, bf16 in MFMA.
// op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k];  op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n].
// Block 256 threads = 2x2 waves, block tile 64x64; each wave one 32x32 MFMA
// tile.  K is staged 128 at a time: each thread fetches its 8 + 8 float4 of
// the next stage into registers while the waves run the current stage's 8
// MFMAs each, so a product with K <= 128 costs one memory round trip and
// K = 256 two (the previous design staged 32 at a time, one round trip each,
// with nothing in flight during the MFMAs).  LDS holds the stage as bf16
// [m][k] and [n][k], so every lane's 8 consecutive-k operands are one 16-byte
// LDS read.  The float4 loads run along each operand's contiguous dimension;
// the feature dimensions (and lda, ldb) must be multiples of 4, the batch
// dimension (M when !TA, K when TA) may be anything.
//
// MFMA 32x32x16 bf16 lane layout: A lane l holds A[l%32][8*(l/32) + i],
// B lane l holds B[8*(l/32) + i][l%32], D lane l element j holds
// D[8*(j/4) + 4*(l/32) + (j%4)][l%32].

enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2 };

// Split-K: block z covers k in [z * kchunk, min(K, (z + 1) * kchunk)) and
// writes its partial product to C + z * M * ldc (kchunk = K: one split).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
                                              const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                              const float* __restrict__ bias, float* __restrict__ Zout,
                                              int kchunk) {
  constexpr int BM = 64, BN = 64, BK = 128, LP = BK + 8;  // +8 bf16 pad per row
  constexpr int NV = (BM * BK / 4) / 256;                 // float4 per thread per operand and stage (8)
  static_assert(BM == BN && NV == 8, "tile shape");
  __shared__ uint16_t As[BM][LP];
  __shared__ uint16_t Bs[BN][LP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  C += (size_t)blockIdx.z * M * ldc;
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 ra[NV], rb[NV];
  // Element (i, k) of a stage: A rows i = m, B rows i = n.  A float4 covers
  // 4 consecutive k (row-major in k) or 4 consecutive rows (k-major).
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int idx = tid + 256 * e;
      if (!TA) {  // A[m][k]: 4 consecutive k
        const int mm = idx / (BK / 4), gk = k0 + 4 * (idx % (BK / 4)), gm = m0 + mm;
        ra[e] = (gm < M && gk < k_hi) ? *reinterpret_cast<const float4*>(A + (size_t)gm * lda + gk) : zero;
      } else {    // A[k][m]: 4 consecutive m
        const int gk = k0 + idx / (BM / 4), gm = m0 + 4 * (idx % (BM / 4));
        ra[e] = (gk < k_hi && gm < M) ? *reinterpret_cast<const float4*>(A + (size_t)gk * lda + gm) : zero;
      }
      if (TB) {   // B[n][k]: 4 consecutive k
        const int nn = idx / (BK / 4), gk = k0 + 4 * (idx % (BK / 4)), gn = n0 + nn;
        rb[e] = (gn < N && gk < k_hi) ? *reinterpret_cast<const float4*>(B + (size_t)gn * ldb + gk) : zero;
      } else {    // B[k][n]: 4 consecutive n
        const int gk = k0 + idx / (BN / 4), gn = n0 + 4 * (idx % (BN / 4));
        rb[e] = (gk < k_hi && gn < N) ? *reinterpret_cast<const float4*>(B + (size_t)gk * ldb + gn) : zero;
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int idx = tid + 256 * e;
      const float va[4] = {ra[e].x, ra[e].y, ra[e].z, ra[e].w};
      const float vb[4] = {rb[e].x, rb[e].y, rb[e].z, rb[e].w};
      if (!TA) {
        const int mm = idx / (BK / 4), kk = 4 * (idx % (BK / 4));
        const uint2 w = make_uint2((uint32_t)to_bf16_bits(va[0]) | ((uint32_t)to_bf16_bits(va[1]) << 16),
                                   (uint32_t)to_bf16_bits(va[2]) | ((uint32_t)to_bf16_bits(va[3]) << 16));
        *reinterpret_cast<uint2*>(&As[mm][kk]) = w;
      } else {
        const int kk = idx / (BM / 4), mm = 4 * (idx % (BM / 4));
#pragma unroll
        for (int j = 0; j < 4; ++j) As[mm + j][kk] = to_bf16_bits(va[j]);
      }
      if (TB) {
        const int nn = idx / (BK / 4), kk = 4 * (idx % (BK / 4));
        const uint2 w = make_uint2((uint32_t)to_bf16_bits(vb[0]) | ((uint32_t)to_bf16_bits(vb[1]) << 16),
                                   (uint32_t)to_bf16_bits(vb[2]) | ((uint32_t)to_bf16_bits(vb[3]) << 16));
        *reinterpret_cast<uint2*>(&Bs[nn][kk]) = w;
      } else {
        const int kk = idx / (BN / 4), nn = 4 * (idx % (BN / 4));
#pragma unroll
        for (int j = 0; j < 4; ++j) Bs[nn + j][kk] = to_bf16_bits(vb[j]);
      }
    }
  };
  f32x16 acc = {};
  if (k_lo < k_hi) load(k_lo);
  for (int k0 = k_lo; k0 < k_hi; k0 += BK) {
    __syncthreads();  // the previous stage's MFMAs have read LDS
    stash();
    __syncthreads();
    if (k0 + BK < k_hi) load(k0 + BK);  // in flight during this stage's MFMAs
    const int kn = min(BK, k_hi - k0);   // zero-filled beyond k_hi: whole 16-k steps only
    const int half = lane >> 5, r = lane & 31;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      if (kk < kn) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + r][kk + 8 * half]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + r][kk + 8 * half]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
  }
  const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = m0 + wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    if (row < M && col < N) {
      float v = acc[j];
      if (EPI != EPI_STORE) v += bias[col];
      if (EPI == EPI_BIAS_SILU) {
        Zout[(size_t)row * ldc + col] = v;
        v = silu(v);
      }
      C[(size_t)row * ldc + col] = v;
    }
  }
}

// The float4 loads need the feature dimensions and leading dimensions to be
// multiples of 4 (and the operands 16-byte aligned).
template <bool TA, bool TB>
bool gemm_shape_ok(int M, int N, int K, const float* A, int lda, const float* B, int ldb) {
  const bool a_ok = (TA ? M : K) % 4 == 0 && lda % 4 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
  const bool b_ok = (TB ? K : N) % 4 == 0 && ldb % 4 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0;
  return a_ok && b_ok;
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, const float* bias = nullptr, float* Z = nullptr) {
  if (!gemm_shape_ok<TA, TB>(M, N, K, A, lda, B, ldb)) return hipErrorInvalidValue;
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias, Z, K);
  return hipGetLastError();
}

// Finishing kernel of a weight gradient: threads [0, mn) sum the split-K
// partials of C[M x N] (ldc = N) in split order (none when splits == 0: the
// product went straight to C); threads [mn, mn + nb) sum the bias gradient's
// per-block column partials (k_dz_partials) in block order.  The loads of a
// thread are issued eight at a time ahead of the adds, which keep their order.
__global__ void k_finish_grads(size_t mn, int splits, const float* __restrict__ ws, float* __restrict__ C, int nb,
                               int parts, const float* __restrict__ part, float* __restrict__ db) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  auto sum = [](const float* p, size_t stride, int count) {
    float acc = 0.0f;
    int z = 0;
    for (; z + 8 <= count; z += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(z + u) * stride];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = (z + u == 0) ? v[u] : acc + v[u];
    }
    for (; z < count; ++z) acc = z == 0 ? p[(size_t)z * stride] : acc + p[(size_t)z * stride];
    return acc;
  };
  if (i < mn) {
    if (splits > 0) C[i] = sum(ws + i, mn, splits);
  } else if (i < mn + (size_t)nb) {
    const int n = (int)(i - mn);
    db[n] = sum(part + n, (size_t)nb, parts);
  }
}

constexpr int kMaxSplits = 64;
constexpr int kDzRows = 64;  // rows per part of k_dz_partials
static_assert(kDzRows % 4 == 0, "four waves share a part");

// Weight gradients have a small output (<= 256 x 256) and K = the batch: the
// K range is split over enough blocks to fill the chip, partials summed in a
// fixed order (deterministic).  The bias gradient's column partials (nb
// columns, `parts` blocks of k_dz_partials) are summed in the same finishing
// launch.
template <bool TA, bool TB>
hipError_t gemm_splitk(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                       float* C, float* ws, int nb, int parts, const float* part, float* db) {
  const int tiles = ((N + 63) / 64) * ((M + 63) / 64);
  int splits = std::min(kMaxSplits, std::max(1, 1024 / tiles));
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  splits = (K + kchunk - 1) / kchunk;
  hipError_t e;
  if (splits <= 1) {
    if ((e = gemm<TA, TB, EPI_STORE>(st, M, N, K, A, lda, B, ldb, C, N)) != hipSuccess) return e;
    splits = 0;
  } else {
    if (!gemm_shape_ok<TA, TB>(M, N, K, A, lda, B, ldb)) return hipErrorInvalidValue;
    dim3 grid((N + 63) / 64, (M + 63) / 64, splits);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI_STORE>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, ws, N, nullptr,
                       nullptr, kchunk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const size_t mn = (size_t)M * N, total = mn + (size_t)nb;
  hipLaunchKernelGGL(k_finish_grads, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, mn, splits, ws, C, nb,
                     parts, part, db);
  return hipGetLastError();
}

// Column partials of dZ over parts of kDzRows rows.  A 256-thread block
// takes 64 columns of one part, each of its four waves a quarter of the rows:
// with SILU it writes dZ = dH * silu'(Z) (without it dZ is dH itself, the
// output layer), every value also into LDS; then wave 0 sums each column's
// values in row order, so part[b][n] = sum of the part's rows in order.
template <bool SILU>
__global__ void __launch_bounds__(256) k_dz_partials(int M, int N, const float* __restrict__ dH,
                                                     const float* __restrict__ Z, float* __restrict__ dZ,
                                                     float* __restrict__ part) {
  constexpr int RPW = kDzRows / 4;  // rows per wave
  __shared__ float sv[kDzRows][65];
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = blockIdx.y * 64 + c;
  const int r0 = blockIdx.x * kDzRows;
  if (n < N) {
    float v[RPW], z[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
      const int r = r0 + w * RPW + u;
      v[u] = r < M ? dH[(size_t)r * N + n] : 0.0f;
      if (SILU) z[u] = r < M ? Z[(size_t)r * N + n] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
      const int r = r0 + w * RPW + u;
      float x = v[u];
      if (SILU) {
        x *= silu_grad(z[u]);
        if (r < M) dZ[(size_t)r * N + n] = x;
      }
      sv[w * RPW + u][c] = x;
    }
  }
  __syncthreads();
  if (w == 0 && n < N) {
    const int rows = min(kDzRows, M - r0);
    float acc = 0.0f;
    for (int u = 0; u < rows; ++u) acc = u == 0 ? sv[u][c] : acc + sv[u][c];
    part[(size_t)blockIdx.x * N + n] = acc;
  }
}

// MSE on normalised targets: dY = 2 (Y - T) / (M * O), loss partials per block
__global__ void __launch_bounds__(256) k_mse(size_t n, const float* __restrict__ Y, const float* __restrict__ T,
                                             float inv, float* __restrict__ dY, float* __restrict__ partial) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < n) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * inv;
    e2 = d * d;
  }
  __shared__ float red[256];
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba) on the flat FP32 parameter vector
__global__ void k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// Fused inference: 128 cells per 256-thread block, the activations in one LDS
// buffer as bf16 [128 cells][K + 8] (66 KB: two blocks per CU, so that one
// block's epilogue overlaps the other's MFMAs), overwritten layer by layer.  Each layer is computed transposed, D^T = W^T H^T: the
// packed weight fragments are the MFMA A operand and the activation rows the
// B operand, so an accumulator lane holds four consecutive features of one
// cell per register group and the epilogue writes them as one 8-byte LDS
// store (and, in the last layer, one 16-byte global store).  Wave w computes
// all 128 cells of a hidden layer for features [64 w, +64) as 4 x 2 MFMA
// tiles of 32x32, so each weight fragment feeds four MFMAs and each
// activation fragment two, and the weights are read once per 128 cells.
// packed[((nt * KS) + ks) * 64 + l] holds lane l's 8 k-values of a (k-step,
// feature tile) weight fragment for KS = K/16 k-steps; weight fragments are
// loaded two k-steps and activation fragments one k-step ahead of the MFMAs.
// SiLU uses the hardware reciprocal: the result is rounded to bf16 anyway.
constexpr int kFM = 128;           // cells per block
constexpr int kInferThreads = 256;
constexpr int kLDA = kHidden + 8;  // LDS row pitch (bf16)

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
static_assert(kOut % 4 == 0 && kHidden == 64 * (kInferThreads / 64) && kFM == 128 && kIn % 16 == 0, "inference tiling");

__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// accumulator lane l, register j of a transposed 32x32 tile: cell l % 32,
// feature 8 (j / 4) + 4 (l / 32) + j % 4
__device__ __forceinline__ void store_tile_act(uint16_t* act, int cell0, int feat0, const f32x16& acc,
                                               const float* __restrict__ bias, int lane) {
  const int cell = cell0 + (lane & 31);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int f = feat0 + 8 * g + 4 * (lane >> 5);
    const float4 bc = *reinterpret_cast<const float4*>(bias + f);
    bf16x4 v;
    v.x = (__bf16)silu_fast(acc[4 * g + 0] + bc.x);
    v.y = (__bf16)silu_fast(acc[4 * g + 1] + bc.y);
    v.z = (__bf16)silu_fast(acc[4 * g + 2] + bc.z);
    v.w = (__bf16)silu_fast(acc[4 * g + 3] + bc.w);
    *reinterpret_cast<bf16x4*>(&act[cell * kLDA + f]) = v;
  }
}

template <int KS>
__device__ __forceinline__ void hidden_layer(uint16_t* act, const bf16x8* __restrict__ Wp,
                                             const float* __restrict__ bias, int wave, int lane) {
  f32x16 acc[4][2] = {};
  const int half = lane >> 5, r = lane & 31;
  const int cell0 = 0, nt0 = 2 * wave;
  const bf16x8* w0 = Wp + (size_t)nt0 * KS * 64 + lane;
  const bf16x8* w1 = w0 + (size_t)KS * 64;
  const uint16_t* hrow = act + (cell0 + r) * kLDA + 8 * half;
  // weight fragments two k-steps ahead (a ring of three), activation
  // fragments one k-step ahead; the loop is unrolled, so ring slots are
  // registers
  bf16x8 wa[3][2], h[2][4];
  wa[0][0] = w0[0];
  wa[0][1] = w1[0];
  if (KS > 1) {
    wa[1][0] = w0[64];
    wa[1][1] = w1[64];
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) h[0][mt] = *reinterpret_cast<const bf16x8*>(hrow + mt * 32 * kLDA);
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    if (ks + 2 < KS) {
      wa[(ks + 2) % 3][0] = w0[(ks + 2) * 64];
      wa[(ks + 2) % 3][1] = w1[(ks + 2) * 64];
    }
    if (ks + 1 < KS) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        h[(ks + 1) & 1][mt] = *reinterpret_cast<const bf16x8*>(hrow + mt * 32 * kLDA + (ks + 1) * 16);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      acc[mt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[ks % 3][0], h[ks & 1][mt], acc[mt][0], 0, 0, 0);
      acc[mt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[ks % 3][1], h[ks & 1][mt], acc[mt][1], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read this layer's input
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) store_tile_act(act, cell0 + mt * 32, (nt0 + nt) * 32, acc[mt][nt], bias, lane);
  __syncthreads();
}

__global__ void __launch_bounds__(kInferThreads) k_mlp_infer(int M, const uint16_t* __restrict__ X /* [M][kIn] bf16 */,
                                                             const bf16x8* __restrict__ W0p,
                                                             const bf16x8* __restrict__ W1p,
                                                             const bf16x8* __restrict__ W2p,
                                                             const bf16x8* __restrict__ W3p,
                                                             const float* __restrict__ b0, const float* __restrict__ b1,
                                                             const float* __restrict__ b2, const float* __restrict__ b3,
                                                             float* __restrict__ Y /* [M][kOut] */) {
  __shared__ uint16_t act[kFM * kLDA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kFM;
  // the 128 x 80 input tile, 16 bytes per thread and step (zero rows past M);
  // rows are 160 bytes, so every chunk is 16-byte aligned
  constexpr int kChunks = kIn / 8;
  for (int e = tid; e < kFM * kChunks; e += kInferThreads) {
    const int rr = e / kChunks, cc = (e % kChunks) * 8;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (m0 + rr < M) v = *reinterpret_cast<const uint4*>(X + (size_t)(m0 + rr) * kIn + cc);
    *reinterpret_cast<uint4*>(&act[rr * kLDA + cc]) = v;
  }
  __syncthreads();
  hidden_layer<kIn / 16>(act, W0p, b0, wave, lane);
  hidden_layer<kHidden / 16>(act, W1p, b1, wave, lane);
  hidden_layer<kHidden / 16>(act, W2p, b2, wave, lane);
  // output layer, transposed like the others: wave w takes cells [32w, 32w + 32)
  f32x16 acc = {};
  const int half = lane >> 5, r = lane & 31;
#pragma unroll 4
  for (int ks = 0; ks < kHidden / 16; ++ks) {
    const bf16x8 h = *reinterpret_cast<const bf16x8*>(&act[(wave * 32 + r) * kLDA + ks * 16 + 8 * half]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W3p[(size_t)ks * 64 + lane], h, acc, 0, 0, 0);
  }
  const int cell = m0 + wave * 32 + r;
  if (cell < M) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int f = 8 * g + 4 * half;  // features f..f+3 (kOut = 28 is a multiple of 4)
      if (f < kOut) {
        const float4 bc = *reinterpret_cast<const float4*>(b3 + f);
        *reinterpret_cast<float4*>(Y + (size_t)cell * kOut + f) =
            make_float4(acc[4 * g] + bc.x, acc[4 * g + 1] + bc.y, acc[4 * g + 2] + bc.z, acc[4 * g + 3] + bc.w);
      }
    }
  }
}

// Packs a row-major f32 weight [K x N] into the per-lane bf16 fragment image
// for N padded to a multiple of 32 (zero columns past N).
__global__ void k_pack_weights(const float* __restrict__ W, int K, int N, int NT, uint16_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one lane-fragment per thread
  const int KS = K / 16;
  if (t >= NT * KS * 64) return;
  const int l = t % 64, ks = (t / 64) % KS, nt = t / (64 * KS);
  const int n = nt * 32 + (l & 31);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = ks * 16 + 8 * (l >> 5) + i;
    const float v = (n < N) ? W[(size_t)k * N + n] : 0.0f;
    out[(size_t)t * 8 + i] = to_bf16_bits(v);
  }
}

// normalised f32 features -> bf16
__global__ void k_to_bf16(size_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16_bits(x[i]);
}

inline unsigned g1(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// MFMA calibration: each wave runs kMfmaChains independent chains of
// v_mfma_f32_32x32x16_bf16 on register operands for `iters` steps, with no
// memory traffic but one store per lane at the end (which keeps the products
// live): the dense BF16 rate the chip sustains, the denominator measured next
// to the vendor's.
constexpr int kMfmaChains = 4;
constexpr int kMfmaProbeIters = 4096;
constexpr int kMfmaProbeBlocksPerCu = 4;  // 256-thread blocks: 4 waves per SIMD, all resident at once

__global__ void __launch_bounds__(256) k_mfma_probe(int iters, float seed, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  bf16x8 a, b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = (__bf16)(seed * (float)(lane + i + 1));
    b[i] = (__bf16)(seed * (float)(64 - lane + i));
  }
  f32x16 acc[kMfmaChains];
#pragma unroll
  for (int c = 0; c < kMfmaChains; ++c) acc[c] = f32x16{};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < kMfmaChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
  }
  float sum = 0.0f;
#pragma unroll
  for (int c = 0; c < kMfmaChains; ++c)
#pragma unroll
    for (int j = 0; j < 16; ++j) sum += acc[c][j];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

MlpLayout mlp_layout() {
  MlpLayout L;
  const int dims[5] = {kIn, kHidden, kHidden, kHidden, kOut};
  size_t off = 0;
  for (int l = 0; l < 4; ++l) {
    L.in[l] = dims[l];
    L.out[l] = dims[l + 1];
    L.w[l] = off;
    off += (size_t)dims[l] * dims[l + 1];
    L.b[l] = off;
    off += dims[l + 1];
  }
  L.total = off;
  return L;
}

int mlp_alloc(marl_ctx* ctx, MlpState& s, int batch) {
  const MlpLayout L = mlp_layout();
  if (!s.w) {
    MARL_HIP(ctx, hipMalloc(&s.w, L.total * sizeof(float)));
    MARL_HIP(ctx, hipMalloc(&s.g, L.total * sizeof(float)));
    MARL_HIP(ctx, hipMalloc(&s.m, L.total * sizeof(float)));
    MARL_HIP(ctx, hipMalloc(&s.v, L.total * sizeof(float)));
    MARL_HIP(ctx, hipMemsetAsync(s.m, 0, L.total * sizeof(float), ctx->stream));
    MARL_HIP(ctx, hipMemsetAsync(s.v, 0, L.total * sizeof(float), ctx->stream));
    // packed images: layers 0-2 have 256 outputs (8 n-tiles), layer 3 one n-tile
    const size_t p0 = (size_t)8 * (kIn / 16) * 64 * 8, p12 = (size_t)8 * (kHidden / 16) * 64 * 8,
                 p3 = (size_t)1 * (kHidden / 16) * 64 * 8;
    MARL_HIP(ctx, hipMalloc(&s.packed, (p0 + 2 * p12 + p3) * sizeof(uint16_t)));
    s.pk[0] = s.packed;
    s.pk[1] = s.pk[0] + p0;
    s.pk[2] = s.pk[1] + p12;
    s.pk[3] = s.pk[2] + p12;
    MARL_HIP(ctx, hipMalloc(&s.norm, sizeof(float) * 2 * (kIn + kOut)));
    MARL_HIP(ctx, hipMalloc(&s.ws, sizeof(float) * kMaxSplits * kHidden * kHidden));
    s.step = 0;
  }
  if (batch > s.batch) {
    float** bufs[] = {&s.x, &s.t, &s.y, &s.dy};
    for (float** b : bufs)
      if (*b) (void)hipFree(*b);
    for (int l = 0; l < 3; ++l) {
      if (s.z[l]) (void)hipFree(s.z[l]);
      if (s.h[l]) (void)hipFree(s.h[l]);
    }
    if (s.dh) (void)hipFree(s.dh);
    if (s.dz) (void)hipFree(s.dz);
    if (s.loss_part) (void)hipFree(s.loss_part);
    if (s.dzpart) (void)hipFree(s.dzpart);
    MARL_HIP(ctx, hipMalloc(&s.x, sizeof(float) * batch * kIn));
    MARL_HIP(ctx, hipMalloc(&s.t, sizeof(float) * batch * kOut));
    MARL_HIP(ctx, hipMalloc(&s.y, sizeof(float) * batch * kOut));
    MARL_HIP(ctx, hipMalloc(&s.dy, sizeof(float) * batch * kOut));
    for (int l = 0; l < 3; ++l) {
      MARL_HIP(ctx, hipMalloc(&s.z[l], sizeof(float) * batch * kHidden));
      MARL_HIP(ctx, hipMalloc(&s.h[l], sizeof(float) * batch * kHidden));
    }
    MARL_HIP(ctx, hipMalloc(&s.dh, sizeof(float) * batch * kHidden));
    MARL_HIP(ctx, hipMalloc(&s.dz, sizeof(float) * batch * kHidden));
    MARL_HIP(ctx, hipMalloc(&s.loss_part, sizeof(float) * (g1((size_t)batch * kOut) + 1)));
    MARL_HIP(ctx, hipMalloc(&s.dzpart, sizeof(float) * ((batch + kDzRows - 1) / kDzRows) * kHidden));
    s.batch = batch;
  }
  return MARL_OK;
}

void mlp_free(MlpState& s) {
  void* bufs[] = {s.w, s.g, s.m, s.v, s.packed, s.norm, s.ws, s.x, s.t, s.y, s.dy, s.dh, s.dz, s.loss_part, s.dzpart,
                  s.z[0], s.z[1], s.z[2], s.h[0], s.h[1], s.h[2], s.xb, s.yinf};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  s = MlpState{};
}

// forward on s.x[0..M) (normalised features) -> s.y, keeping Z/H for backward
int mlp_forward_train(marl_ctx* ctx, MlpState& s, int M) {
  const MlpLayout L = mlp_layout();
  const float* in = s.x;
  int K = kIn;
  for (int l = 0; l < 3; ++l) {
    MARL_HIP(ctx, (gemm<false, false, EPI_BIAS_SILU>(ctx->stream, M, kHidden, K, in, K, s.w + L.w[l], kHidden, s.h[l],
                                                      kHidden, s.w + L.b[l], s.z[l])));
    in = s.h[l];
    K = kHidden;
  }
  MARL_HIP(ctx, (gemm<false, false, EPI_BIAS>(ctx->stream, M, kOut, kHidden, in, kHidden, s.w + L.w[3], kOut, s.y,
                                              kOut, s.w + L.b[3], nullptr)));
  return MARL_OK;
}

// backward from s.dy (dL/dY) into s.g
int mlp_backward(marl_ctx* ctx, MlpState& s, int M) {
  const MlpLayout L = mlp_layout();
  const int parts = (M + kDzRows - 1) / kDzRows;
  // layer 3: dY's column partials, then dW4 with db4 in its finishing launch
  hipLaunchKernelGGL(k_dz_partials<false>, dim3(parts, (kOut + 63) / 64), dim3(256), 0, ctx->stream, M, kOut, s.dy,
                     nullptr, nullptr, s.dzpart);
  MARL_HIP(ctx, hipGetLastError());
  MARL_HIP(ctx, (gemm_splitk<true, false>(ctx->stream, kHidden, kOut, M, s.h[2], kHidden, s.dy, kOut,
                                          s.g + L.w[3], s.ws, kOut, parts, s.dzpart, s.g + L.b[3])));
  MARL_HIP(ctx, (gemm<false, true, EPI_STORE>(ctx->stream, M, kHidden, kOut, s.dy, kOut, s.w + L.w[3], kOut, s.dh,
                                              kHidden)));
  for (int l = 2; l >= 0; --l) {
    // dZ = dH * silu'(Z) and its column partials in one pass
    hipLaunchKernelGGL(k_dz_partials<true>, dim3(parts, (kHidden + 63) / 64), dim3(256), 0, ctx->stream, M, kHidden,
                       s.dh, s.z[l], s.dz, s.dzpart);
    MARL_HIP(ctx, hipGetLastError());
    const float* hin = l > 0 ? s.h[l - 1] : s.x;
    const int K = l > 0 ? kHidden : kIn;
    MARL_HIP(ctx, (gemm_splitk<true, false>(ctx->stream, K, kHidden, M, hin, K, s.dz, kHidden, s.g + L.w[l], s.ws,
                                            kHidden, parts, s.dzpart, s.g + L.b[l])));
    if (l > 0)
      MARL_HIP(ctx, (gemm<false, true, EPI_STORE>(ctx->stream, M, K, kHidden, s.dz, kHidden, s.w + L.w[l], kHidden,
                                                  s.dh, kHidden)));
  }
  return MARL_OK;
}

int mlp_loss(marl_ctx* ctx, MlpState& s, int M, float* loss_host) {
  const size_t n = (size_t)M * kOut;
  const unsigned nb = g1(n);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, ctx->stream, n, s.y, s.t, 1.0f / (float)n, s.dy, s.loss_part);
  MARL_HIP(ctx, hipGetLastError());
  if (loss_host) {
    std::vector<float> part(nb);
    MARL_HIP(ctx, hipMemcpyAsync(part.data(), s.loss_part, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double acc = 0.0;
    for (float p : part) acc += p;
    *loss_host = (float)(acc / (double)n);
  }
  return MARL_OK;
}

int mlp_adam(marl_ctx* ctx, MlpState& s, float lr) {
  const MlpLayout L = mlp_layout();
  s.step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)s.step), c2 = 1.0f - powf(b2, (float)s.step);
  hipLaunchKernelGGL(k_adam, dim3(g1(L.total)), dim3(256), 0, ctx->stream, L.total, s.w, s.g, s.m, s.v, lr, b1, b2,
                     eps, c1, c2);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

int mlp_pack(marl_ctx* ctx, MlpState& s) {
  const MlpLayout L = mlp_layout();
  const int NT[4] = {8, 8, 8, 1};
  for (int l = 0; l < 4; ++l) {
    const int frags = NT[l] * (L.in[l] / 16) * 64;
    hipLaunchKernelGGL(k_pack_weights, dim3(g1(frags)), dim3(256), 0, ctx->stream, s.w + L.w[l], L.in[l], L.out[l],
                       NT[l], s.pk[l]);
    MARL_HIP(ctx, hipGetLastError());
  }
  return MARL_OK;
}

// fused inference on normalised f32 features xin[M][kIn] (device) -> yout[M][kOut]
int mlp_infer(marl_ctx* ctx, MlpState& s, int M, const float* xin, float* yout) {
  if (M <= 0) return MARL_OK;
  if ((size_t)M > s.infer_cap) {
    if (s.xb) (void)hipFree(s.xb);
    s.xb = nullptr;
    MARL_HIP(ctx, hipMalloc(&s.xb, sizeof(uint16_t) * (size_t)M * kIn));
    s.infer_cap = (size_t)M;
  }
  hipLaunchKernelGGL(k_to_bf16, dim3(g1((size_t)M * kIn)), dim3(256), 0, ctx->stream, (size_t)M * kIn, xin, s.xb);
  MARL_HIP(ctx, hipGetLastError());
  return mlp_infer_bf16(ctx, s, M, s.xb, yout);
}

int mlp_infer_bf16(marl_ctx* ctx, MlpState& s, int M, const uint16_t* xb, float* yout) {
  if (M <= 0) return MARL_OK;
  const MlpLayout L = mlp_layout();
  // the fused kernel moves inputs, biases and outputs in 16-byte vectors
  auto misaligned = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) != 0; };
  bool bad = misaligned(xb) || misaligned(yout);
  for (int l = 0; l < 4; ++l) bad = bad || misaligned(s.w + L.b[l]);
  if (bad) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "surrogate inference buffers must be 16-byte aligned");
  hipLaunchKernelGGL(k_mlp_infer, dim3(g1(M, kFM)), dim3(kInferThreads), 0, ctx->stream, M, xb,
                     reinterpret_cast<const bf16x8*>(s.pk[0]), reinterpret_cast<const bf16x8*>(s.pk[1]),
                     reinterpret_cast<const bf16x8*>(s.pk[2]), reinterpret_cast<const bf16x8*>(s.pk[3]),
                     s.w + L.b[0], s.w + L.b[1], s.w + L.b[2], s.w + L.b[3], yout);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// out: kMfmaProbeBlocksPerCu * cus * 256 floats.  *flop = FLOP per launch.
int marl_launch_mfma_probe(marl_ctx* ctx, int cus, float* out, double* flop) {
  const int blocks = kMfmaProbeBlocksPerCu * cus;
  hipLaunchKernelGGL(k_mfma_probe, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, kMfmaProbeIters, 1.0f / 64.0f,
                     out);
  MARL_HIP(ctx, hipGetLastError());
  if (flop) *flop = (double)blocks * 4.0 * kMfmaProbeIters * kMfmaChains * (2.0 * 32 * 32 * 16);
  return MARL_OK;
}

size_t marl_mfma_probe_floats(int cus) { return (size_t)kMfmaProbeBlocksPerCu * cus * 256; }
