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
// Block 256 threads = 2x2 waves, block tile 64x64, K tile 32; each wave one
// 32x32 MFMA tile.  LDS holds the A tile as [m][k] and the B tile as [n][k]
// so that every lane's 8 consecutive-k operands are one 16-byte LDS read.
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
  constexpr int BM = 64, BN = 64, BK = 32, LP = BK + 8;  // +8 bf16 pad per row
  __shared__ uint16_t As[BM][LP];
  __shared__ uint16_t Bs[BN][LP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  C += (size_t)blockIdx.z * M * ldc;
  f32x16 acc = {};
  for (int k0 = k_lo; k0 < k_hi; k0 += BK) {
    // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + 256 * e;
      int mm, kk;
      if (TA) { kk = idx / BM; mm = idx % BM; }   // walk m fastest: A[k][m] rows are contiguous
      else { mm = idx / BK; kk = idx % BK; }
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.0f;
      if (gm < M && gk < k_hi) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[mm][kk] = to_bf16_bits(v);
      int nn, kb;
      if (TB) { nn = idx / BK; kb = idx % BK; }   // B[n][k] rows contiguous in k
      else { kb = idx / BN; nn = idx % BN; }
      const int gn = n0 + nn, gkb = k0 + kb;
      float w = 0.0f;
      if (gn < N && gkb < k_hi) w = TB ? B[(size_t)gn * ldb + gkb] : B[(size_t)gkb * ldb + gn];
      Bs[nn][kb] = to_bf16_bits(w);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int half = lane >> 5, r = lane & 31;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + r][kk + 8 * half]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + r][kk + 8 * half]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
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

template <bool TA, bool TB, int EPI>
hipError_t gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, const float* bias = nullptr, float* Z = nullptr) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias, Z, K);
  return hipGetLastError();
}

// C[M x N] (ldc = N) = sum over splits of the partials in ws, in split order
__global__ void k_sum_splits(size_t mn, int splits, const float* __restrict__ ws, float* __restrict__ C) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float acc = ws[i];
  for (int z = 1; z < splits; ++z) acc += ws[(size_t)z * mn + i];
  C[i] = acc;
}

constexpr int kMaxSplits = 64;

// Weight gradients have a small output (<= 256 x 256) and K = the batch: the
// K range is split over enough blocks to fill the chip, partials summed in a
// fixed order (deterministic).
template <bool TA, bool TB>
hipError_t gemm_splitk(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                       float* C, float* ws) {
  const int tiles = ((N + 63) / 64) * ((M + 63) / 64);
  int splits = std::min(kMaxSplits, std::max(1, 1024 / tiles));
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  splits = (K + kchunk - 1) / kchunk;
  if (splits <= 1) return gemm<TA, TB, EPI_STORE>(st, M, N, K, A, lda, B, ldb, C, N);
  dim3 grid((N + 63) / 64, (M + 63) / 64, splits);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI_STORE>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, ws, N, nullptr,
                     nullptr, kchunk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t mn = (size_t)M * N;
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, mn, splits, ws, C);
  return hipGetLastError();
}

// dZ = dH * silu'(Z), elementwise
__global__ void k_silu_back(size_t n, const float* __restrict__ dH, const float* __restrict__ Z,
                            float* __restrict__ dZ) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = dH[i] * silu_grad(Z[i]);
}

// db[n] = sum_m dZ[m][n]: one block per column group, fixed-order reduction
__global__ void __launch_bounds__(256) k_bias_grad(int M, int N, const float* __restrict__ dZ,
                                                   float* __restrict__ db) {
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int m = threadIdx.x; m < M; m += 256) s += dZ[(size_t)m * N + n];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = red[0];
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
// Fused inference: 64 cells per 256-thread block, activations in LDS as bf16
// [64][K + 8].  Wave w computes output columns [64w, 64w + 64) of every
// hidden layer (2 x 2 MFMA tiles of 32x32), reading B operands from a packed
// weight image in which lane l's 8 k-values of a (k-step, n-tile) fragment are
// one 16-byte word: packed[((nt * KS) + ks) * 64 + l] for layer weights of
// KS = K/16 k-steps.  The last layer (28 outputs, one 32-wide n-tile) is
// computed by waves 0-1, one 32-row m-tile each.
constexpr int kFM = 64;          // cells per block
constexpr int kLDA = kHidden + 8;  // LDS row pitch (bf16)

__device__ __forceinline__ void store_tile_act(uint16_t* out, int row0, int col0, const f32x16& acc,
                                               const float* __restrict__ bias, int lane) {
  const int col = col0 + (lane & 31);
  const float bc = bias[col];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = row0 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    out[row * kLDA + col] = to_bf16_bits(silu(acc[j] + bc));
  }
}

template <int KS>
__device__ __forceinline__ void hidden_layer(const uint16_t* in, uint16_t* out, const bf16x8* __restrict__ Wp,
                                             const float* __restrict__ bias, int wave, int lane) {
  f32x16 acc[2][2] = {};
  const int half = lane >> 5, r = lane & 31;
#pragma unroll 4
  for (int ks = 0; ks < KS; ++ks) {
    bf16x8 a[2], b[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) a[mt] = *reinterpret_cast<const bf16x8*>(&in[(mt * 32 + r) * kLDA + ks * 16 + 8 * half]);
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) b[nt] = Wp[((size_t)(wave * 2 + nt) * KS + ks) * 64 + lane];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
  }
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) store_tile_act(out, mt * 32, (wave * 2 + nt) * 32, acc[mt][nt], bias, lane);
}

__global__ void __launch_bounds__(256) k_mlp_infer(int M, const uint16_t* __restrict__ X /* [M][kIn] bf16 */,
                                                   const bf16x8* __restrict__ W0p, const bf16x8* __restrict__ W1p,
                                                   const bf16x8* __restrict__ W2p, const bf16x8* __restrict__ W3p,
                                                   const float* __restrict__ b0, const float* __restrict__ b1,
                                                   const float* __restrict__ b2, const float* __restrict__ b3,
                                                   float* __restrict__ Y /* [M][kOut] */) {
  __shared__ uint16_t bufA[kFM * kLDA];
  __shared__ uint16_t bufB[kFM * kLDA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kFM;
  // load the 64 x 80 input tile (zero rows past M)
  for (int e = tid; e < kFM * kIn; e += 256) {
    const int rr = e / kIn, cc = e % kIn;
    bufA[rr * kLDA + cc] = (m0 + rr < M) ? X[(size_t)(m0 + rr) * kIn + cc] : (uint16_t)0;
  }
  __syncthreads();
  hidden_layer<kIn / 16>(bufA, bufB, W0p, b0, wave, lane);
  __syncthreads();
  hidden_layer<kHidden / 16>(bufB, bufA, W1p, b1, wave, lane);
  __syncthreads();
  hidden_layer<kHidden / 16>(bufA, bufB, W2p, b2, wave, lane);
  __syncthreads();
  if (wave < 2) {
    f32x16 acc = {};
    const int half = lane >> 5, r = lane & 31;
#pragma unroll 4
    for (int ks = 0; ks < kHidden / 16; ++ks) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&bufB[(wave * 32 + r) * kLDA + ks * 16 + 8 * half]);
      const bf16x8 b = W3p[(size_t)ks * 64 + lane];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int col = lane & 31;
    if (col < kOut) {
      const float bc = b3[col];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int row = m0 + wave * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (row < M) Y[(size_t)row * kOut + col] = acc[j] + bc;
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
    s.batch = batch;
  }
  return MARL_OK;
}

void mlp_free(MlpState& s) {
  void* bufs[] = {s.w, s.g, s.m, s.v, s.packed, s.norm, s.ws, s.x, s.t, s.y, s.dy, s.dh, s.dz, s.loss_part,
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
  // layer 3
  MARL_HIP(ctx, (gemm_splitk<true, false>(ctx->stream, kHidden, kOut, M, s.h[2], kHidden, s.dy, kOut,
                                          s.g + L.w[3], s.ws)));
  hipLaunchKernelGGL(k_bias_grad, dim3(kOut), dim3(256), 0, ctx->stream, M, kOut, s.dy, s.g + L.b[3]);
  MARL_HIP(ctx, (gemm<false, true, EPI_STORE>(ctx->stream, M, kHidden, kOut, s.dy, kOut, s.w + L.w[3], kOut, s.dh,
                                              kHidden)));
  for (int l = 2; l >= 0; --l) {
    const size_t n = (size_t)M * kHidden;
    hipLaunchKernelGGL(k_silu_back, dim3(g1(n)), dim3(256), 0, ctx->stream, n, s.dh, s.z[l], s.dz);
    MARL_HIP(ctx, hipGetLastError());
    const float* hin = l > 0 ? s.h[l - 1] : s.x;
    const int K = l > 0 ? kHidden : kIn;
    MARL_HIP(ctx, (gemm_splitk<true, false>(ctx->stream, K, kHidden, M, hin, K, s.dz, kHidden, s.g + L.w[l],
                                            s.ws)));
    hipLaunchKernelGGL(k_bias_grad, dim3(kHidden), dim3(256), 0, ctx->stream, M, kHidden, s.dz, s.g + L.b[l]);
    MARL_HIP(ctx, hipGetLastError());
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
  hipLaunchKernelGGL(k_mlp_infer, dim3(g1(M, kFM)), dim3(256), 0, ctx->stream, M, xb,
                     reinterpret_cast<const bf16x8*>(s.pk[0]), reinterpret_cast<const bf16x8*>(s.pk[1]),
                     reinterpret_cast<const bf16x8*>(s.pk[2]), reinterpret_cast<const bf16x8*>(s.pk[3]),
                     s.w + L.b[0], s.w + L.b[1], s.w + L.b[2], s.w + L.b[3], yout);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}
