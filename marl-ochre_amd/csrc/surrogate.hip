// surrogate.hip — the learned kinetics surrogate: pair harvest, training on
// the GPU, and the per-level surrogate pass with plain fallback and shadow
// evaluation.
//
// No reference counterpart: this replaces the reaction-phase integration of
// CellState::tick (crates/marl-cell/src/cell.rs:223-329) for cells whose
// topology and features lie inside the training distribution and whose
// predicted energy is not near a fate threshold; every other cell runs the
// plain path, so fate decisions near thresholds are the plain kernel's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "marl_internal.h"
#include "surrogate.h"

namespace {

enum StatSlot {
  ST_CELLS = 0, ST_SURR, ST_FB_TOPO, ST_FB_OOD, ST_FB_MARGIN, ST_SHADOW, ST_WITHIN, ST_FATE_AGREE, ST_MAXERR,
  ST_HIST, ST_SPECIES = ST_HIST + kErrBins, ST_FATE_EXACT = ST_SPECIES + 16, ST_FATE_SURR = ST_FATE_EXACT + 4,
  ST_COUNT = ST_FATE_SURR + 4
};

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// binary feature slots (left unnormalised): quiescent, in-prep
__device__ __host__ inline bool binary_feature(int k) { return k == 15 || k == 29; }

constexpr int kSBlock = 128;

// Statistics are counted per block in LDS and flushed with one global atomic
// per non-zero counter: a global atomic per cell would serialise a million
// updates per tick on a handful of addresses.
__device__ __forceinline__ void stats_init(unsigned int* s) {
  for (int k = threadIdx.x; k < ST_COUNT; k += blockDim.x) s[k] = 0u;
  __syncthreads();
}
__device__ __forceinline__ void stats_flush(const unsigned int* s, unsigned long long* __restrict__ g) {
  __syncthreads();
  for (int k = threadIdx.x; k < ST_COUNT; k += blockDim.x) {
    if (!s[k]) continue;
    if (k == ST_MAXERR) atomicMax(&g[k], (unsigned long long)s[k]);
    else atomicAdd(&g[k], (unsigned long long)s[k]);
  }
}

// Features (standardised, BF16) and the explicit step of a level's cells;
// a cell with a continuous feature beyond zmax standard deviations of the
// training pairs is out of distribution and runs the plain path.
__global__ void __launch_bounds__(kSBlock) k_surr_features(const int32_t* __restrict__ order, int count,
                                                         const marl_cell_state* __restrict__ cells,
                                                         const marl_ruleset* __restrict__ rules,
                                                         const float* __restrict__ field,
                                                         const float* __restrict__ light,
                                                         const uint8_t* __restrict__ nbmask, CellParams p,
                                                         const float* __restrict__ norm,
                                                         const float* __restrict__ inv, float zmax,
                                                         uint16_t* __restrict__ featb, float* __restrict__ expl,
                                                         uint8_t* __restrict__ elig,
                                                         unsigned long long* __restrict__ stats) {
  __shared__ float sI[kSBlock * 17];
  __shared__ float sE[kSBlock * 13];
  __shared__ float sQ[kSBlock * 17];
  __shared__ float sR[kSBlock * 13];
  __shared__ float sG[kSBlock * 17];
  __shared__ float sH[kSBlock * 13];
  __shared__ unsigned int sst[ST_COUNT];
  stats_init(sst);
  const int t = blockIdx.x * kSBlock + threadIdx.x;
  if (t < count) {
    const int i = order[t];
    const marl_cell_state c = load_cell(cells + i);
    float* I0 = sI + threadIdx.x * 17;
    float* E = sE + threadIdx.x * 13;
    float* dI = sQ + threadIdx.x * 17;
    float* dD = sR + threadIdx.x * 13;
#pragma unroll
    for (int k = 0; k < MARL_M_INT; ++k) I0[k] = c.internal[k];
    const uint32_t m = nbmask[i];
    const size_t v = ((size_t)c.pos[2] * p.Y + c.pos[1]) * p.X + c.pos[0];
    gather_env6<false>(field, v, m, p, E);
    const float lt = light[v];
    float* gI = sG + threadIdx.x * 17;
    float* gD = sH + threadIdx.x * 13;
    float frel[MARL_R_MAX];
    PtrArr aQ{dI}, aR{dD}, aG{gI}, aH{gD};
    explicit_step(c, AosRules{&rules[i]}, PtrArr{I0}, PtrArr{E}, aQ, aR, lt, p, frel, aG, aH);
    float f[kFeat];
    make_features(c.internal, c, E, lt, dI, dD, frel, f);
    bool ok = true;
    uint16_t* fo = featb + (size_t)t * kFeat;
#pragma unroll
    for (int k = 0; k < kFeat; ++k) {
      const float z = (f[k] - norm[k]) * inv[k];
      if (!binary_feature(k) && !(fabsf(z) <= zmax)) ok = false;
      fo[k] = bf16_bits(z);
    }
    float* xo = expl + (size_t)t * 2 * kTarg;
    target_scales(gI, gD, xo + kTarg);
#pragma unroll
    for (int k = 0; k < MARL_M_INT; ++k) xo[k] = dI[k];
#pragma unroll
    for (int e = 0; e < MARL_S_EXT; ++e) xo[16 + e] = dD[e];
    elig[t] = ok ? 1 : 0;
    atomicAdd(&sst[ST_CELLS], 1u);
    if (!ok) atomicAdd(&sst[ST_FB_OOD], 1u);
  }
  stats_flush(sst, stats);
}

// Shadow error of a cell: the largest relative error over internals 0-14
// (denominator max(|exact|, 1e-3)) into the histogram, and per species
// whether it is within 1e-3.  Is: surrogate, Ip: exact.
__device__ inline void record_error(const float* Is, const float* Ip, unsigned int* sst) {
  float err = 0.0f;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const float ek = fabsf(Is[k] - Ip[k]) / fmaxf(fabsf(Ip[k]), 1e-3f);
    if (ek <= 1e-3f) atomicAdd(&sst[ST_SPECIES + k], 1u);
    err = fmaxf(err, ek);
  }
  atomicAdd(&sst[ST_SHADOW], 1u);
  if (err <= 1e-3f) atomicAdd(&sst[ST_WITHIN], 1u);
  atomicMax(&sst[ST_MAXERR], __builtin_bit_cast(uint32_t, err));
  int bin = (int)floorf((log10f(fmaxf(err, 1e-9f)) + 9.0f) * 2.0f);
  bin = bin < 0 ? 0 : (bin >= kErrBins ? kErrBins - 1 : bin);
  atomicAdd(&sst[ST_HIST + bin], 1u);
}

// One cell of a surrogate level (see k_surr_apply).  The prediction is the
// explicit step plus the network's residual (de-standardised, times the
// output's target scale), clamped as the exact update clamps
// internals ([0, c_max]); internal 15 is the voxel's light, as the tick sets
// it.  The gathered face records are kept for the scatter, as in the
// plain pass.
__device__ __forceinline__ void apply_cell(int i, int t, marl_cell_state* __restrict__ cells,
                                           const marl_ruleset* __restrict__ rules, float* __restrict__ field,
                                           const float* __restrict__ light, const uint8_t* __restrict__ nbmask,
                                           uint8_t* __restrict__ events, const CellParams& p,
                                           const float* __restrict__ out, const float* __restrict__ expl,
                                           const uint8_t* __restrict__ elig, const float* __restrict__ norm,
                                           int mode, float margin, float dead_zone, uint64_t shadow_threshold,
                                           unsigned int* sst,
                                           float* I, float* E, float* D) {
  marl_cell_state c = load_cell(cells + i);
  const AosRules ar{&rules[i]};
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) I[k] = c.internal[k];
  const uint32_t m = nbmask[i];
  const size_t v = ((size_t)c.pos[2] * p.Y + c.pos[1]) * p.X + c.pos[0];
  float4 keep[18];
  gather_env6_keep<false>(field, v, m, p, E, keep);
  const float lt = light[v];
  const float4 fate = ar.fate();

  const float* tmean = norm + 2 * kFeat;
  const float* tstd = tmean + kTarg;
  float Ip[MARL_M_INT], Dp[MARL_S_EXT];
  bool have = elig[t] != 0, use_s = false;
  if (have) {
    const float* y = out + (size_t)t * kTarg;
    const float* x = expl + (size_t)t * 2 * kTarg;
    const float* g = x + kTarg;
    // the correction in target-scale units, dropped inside the dead zone
    auto corr = [&](int k) {
      const float r = y[k] * tstd[k] + tmean[k];
      return fabsf(r) > dead_zone ? r * g[k] : 0.0f;
    };
#pragma unroll
    for (int k = 0; k < 15; ++k) Ip[k] = clamp_rs(c.internal[k] + (x[k] + corr(k)), 0.0f, p.c_max);
    Ip[15] = lt;
#pragma unroll
    for (int e = 0; e < MARL_S_EXT; ++e) Dp[e] = x[16 + e] + corr(16 + e);
    const float e1 = Ip[0];
    const float thr[3] = {fmaxf(fate.y, p.hard_death_floor), fate.x, fate.z};
    bool near = false;
#pragma unroll
    for (int q = 0; q < 3; ++q) near |= fabsf(e1 - thr[q]) < margin;
    if (near) atomicAdd(&sst[ST_FB_MARGIN], 1u);
    else if (mode == MARL_KINETICS_SURROGATE) use_s = true;
  }
  const bool shadow = have && (mode == MARL_KINETICS_SHADOW ||
                               (use_s && (uint64_t)sample_hash(p.stream, (uint32_t)i) < shadow_threshold));
  if (!use_s || shadow) {
    PtrArr aI{I}, aD{D};
    const PtrArr aE{E};
    cell_kinetics(c, ar, aI, aE, aD, lt, p);
  }
  if (shadow) {
    record_error(Ip, I, sst);
    marl_cell_state cs = c, cp = c;
    const uint8_t es = cell_fate(cs, fate, Ip[0], p), ep = cell_fate(cp, fate, I[0], p);
    atomicAdd(&sst[ST_FATE_SURR + es], 1u);
    atomicAdd(&sst[ST_FATE_EXACT + ep], 1u);
    if (es == ep && cs.prep_remaining == cp.prep_remaining && cs.quiescent == cp.quiescent)
      atomicAdd(&sst[ST_FATE_AGREE], 1u);
  }
  if (use_s) {
    atomicAdd(&sst[ST_SURR], 1u);
#pragma unroll
    for (int k = 0; k < MARL_M_INT; ++k) I[k] = Ip[k];
#pragma unroll
    for (int e = 0; e < MARL_S_EXT; ++e) D[e] = Dp[e];
  }
  const uint8_t ev = cell_fate(c, fate, I[0], p);
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) c.internal[k] = I[k];
  cells[i] = c;
  events[i] = ev;
  scatter_from_kept<false>(field, v, m, p, D, keep);
}

// Applies one level: surrogate result where eligible and not near a fate
// threshold, the plain tick otherwise; shadow cells compare both.
__global__ void __launch_bounds__(kSBlock) k_surr_apply(const int32_t* __restrict__ order, int count,
                                                      marl_cell_state* __restrict__ cells,
                                                      const marl_ruleset* __restrict__ rules,
                                                      float* __restrict__ field, const float* __restrict__ light,
                                                      const uint8_t* __restrict__ nbmask,
                                                      uint8_t* __restrict__ events, CellParams p,
                                                      const float* __restrict__ out, const float* __restrict__ expl,
                                                      const uint8_t* __restrict__ elig,
                                                      const float* __restrict__ norm, int mode, float margin,
                                                      float dead_zone, uint64_t shadow_threshold,
                                                      unsigned long long* __restrict__ stats) {
  __shared__ float sI[kSBlock * 17];
  __shared__ float sE[kSBlock * 13];
  __shared__ float sD[kSBlock * 13];
  __shared__ unsigned int sst[ST_COUNT];
  stats_init(sst);
  const int t = blockIdx.x * kSBlock + threadIdx.x;
  if (t < count) apply_cell(order[t], t, cells, rules, field, light, nbmask, events, p, out, expl, elig, norm, mode,
                            margin, dead_zone, shadow_threshold, sst, sI + threadIdx.x * 17, sE + threadIdx.x * 13,
                            sD + threadIdx.x * 13);
  stats_flush(sst, stats);
}

// ---------------------------------------------------------------------------
// training helpers

// column sums and sums of squares (double) of features (with class slots) and targets
__global__ void __launch_bounds__(256) k_colstats(const int32_t* __restrict__ idx, int n, const float* __restrict__ F,
                                                  const float* __restrict__ T, double* __restrict__ acc) {
  constexpr int C = kFeat + kTarg;
  __shared__ double s1[C], s2[C];
  for (int k = threadIdx.x; k < C; k += blockDim.x) s1[k] = s2[k] = 0.0;
  __syncthreads();
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int row = idx[j];
    float f[kFeat];
    for (int k = 0; k < kFeat; ++k) f[k] = F[(size_t)row * kFeat + k];
    for (int k = 0; k < C; ++k) {
      const double x = k < kFeat ? f[k] : T[(size_t)row * kTarg + (k - kFeat)];
      atomicAdd(&s1[k], x);
      atomicAdd(&s2[k], x * x);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < C; k += blockDim.x) {
    atomicAdd(&acc[k], s1[k]);
    atomicAdd(&acc[C + k], s2[k]);
  }
}

// minibatch: rows drawn from idx by a per-(step, slot) counter hash,
// normalised.  A block takes kGatherRows slots: their rows are drawn first
// (one lane each), then the block copies the rows' features and targets with
// consecutive lanes on consecutive elements.
constexpr int kGatherRows = 32;  // the kernel's block is 256 threads
__global__ void __launch_bounds__(256) k_gather_batch(const int32_t* __restrict__ idx, int n, uint64_t seed,
                                                      uint64_t step, int B, const float* __restrict__ F,
                                                      const float* __restrict__ T, const float* __restrict__ norm,
                                                      const float* __restrict__ inv, float* __restrict__ X,
                                                      float* __restrict__ Yt, int sequential) {
  __shared__ int32_t rows[kGatherRows];
  const int b0 = blockIdx.x * kGatherRows;
  if (threadIdx.x < kGatherRows) {
    const int b = b0 + (int)threadIdx.x;
    int row = 0;
    if (b < B) {
      int j;
      if (sequential) {
        j = b % n;
      } else {
        uint64_t c[4] = {step, (uint64_t)b, 0, 0};
        marl_philox4x64(c, seed, 0x5EEDull);
        j = (int)(c[0] % (uint64_t)n);
      }
      row = idx[j];
    }
    rows[threadIdx.x] = row;
  }
  __syncthreads();
  const float* tm = norm + 2 * kFeat;
  const float* ts = tm + kTarg;
  // every element's load is issued before any is stored
  constexpr int kF = (kGatherRows * kFeat + 255) / 256, kT = (kGatherRows * kTarg + 255) / 256;
  float fv[kF], tv[kT];
#pragma unroll
  for (int u = 0; u < kF; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / kFeat, k = e - r * kFeat;
    fv[u] = (e < kGatherRows * kFeat && b0 + r < B) ? F[(size_t)rows[r] * kFeat + k] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < kT; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / kTarg, k = e - r * kTarg;
    tv[u] = (e < kGatherRows * kTarg && b0 + r < B) ? T[(size_t)rows[r] * kTarg + k] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < kF; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / kFeat, k = e - r * kFeat, b = b0 + r;
    if (e < kGatherRows * kFeat && b < B) X[(size_t)b * kFeat + k] = (fv[u] - norm[k]) * inv[k];
  }
#pragma unroll
  for (int u = 0; u < kT; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / kTarg, k = e - r * kTarg, b = b0 + r;
    if (e < kGatherRows * kTarg && b < B) Yt[(size_t)b * kTarg + k] = ts[k] > 0.0f ? (tv[u] - tm[k]) / ts[k] : 0.0f;
  }
}

inline unsigned gs(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Copies on the context stream (which does not synchronise with the null
// stream) and waits for them, so earlier kernels on the stream are ordered
// before the copy and the copy is complete on return.
hipError_t xfer(marl_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}

int ensure_surr(marl_ctx* ctx) {
  if (!ctx->surr) {
    ctx->surr = new SurrogateState();
    MARL_HIP(ctx, hipMalloc(&ctx->surr->stats, sizeof(unsigned long long) * ST_COUNT));
    MARL_HIP(ctx, hipMemsetAsync(ctx->surr->stats, 0, sizeof(unsigned long long) * ST_COUNT, ctx->stream));
    MARL_HIP(ctx, hipMalloc(&ctx->surr->hcount, sizeof(unsigned long long)));
    MARL_HIP(ctx, hipMemsetAsync(ctx->surr->hcount, 0, sizeof(unsigned long long), ctx->stream));
  }
  return MARL_OK;
}

int ensure_level_scratch(marl_ctx* ctx, int64_t n) {
  SurrogateState* s = ctx->surr;
  if (n <= s->scratch_cap) return MARL_OK;
  if (s->featb) (void)hipFree(s->featb);
  if (s->out) (void)hipFree(s->out);
  if (s->expl) (void)hipFree(s->expl);
  if (s->elig) (void)hipFree(s->elig);
  s->featb = nullptr;
  s->out = nullptr;
  s->expl = nullptr;
  s->elig = nullptr;
  const int64_t cap = std::max<int64_t>(n, 65536);
  MARL_HIP(ctx, hipMalloc(&s->featb, sizeof(uint16_t) * cap * kFeat));
  MARL_HIP(ctx, hipMalloc(&s->out, sizeof(float) * cap * kTarg));
  MARL_HIP(ctx, hipMalloc(&s->expl, sizeof(float) * cap * 2 * kTarg));
  MARL_HIP(ctx, hipMalloc(&s->elig, cap));
  s->scratch_cap = cap;
  return MARL_OK;
}

// The normalisation (feature mean/std, target mean/std) and the feature
// reciprocals the kernels multiply by.
int upload_norm(marl_ctx* ctx, SurrogateState* s, const std::vector<float>& nm) {
  if (!s->inv) MARL_HIP(ctx, hipMalloc(&s->inv, sizeof(float) * kFeat));
  std::vector<float> inv(kFeat);
  for (int k = 0; k < kFeat; ++k) inv[k] = 1.0f / nm[kFeat + k];
  MARL_HIP(ctx, xfer(ctx, s->mlp.norm, nm.data(), sizeof(float) * nm.size(), hipMemcpyHostToDevice));
  MARL_HIP(ctx, xfer(ctx, s->inv, inv.data(), sizeof(float) * kFeat, hipMemcpyHostToDevice));
  return MARL_OK;
}


// deterministic standardised-looking bf16 inputs for timing the network alone
__global__ void k_fill_timing_input(size_t n, uint16_t* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = bf16_bits(((float)((i * 2654435761u) & 1023u) - 512.0f) * (1.0f / 256.0f));
}

}  // namespace

void surrogate_free(SurrogateState* s) {
  if (!s) return;
  mlp_free(s->mlp);
  void* bufs[] = {s->hfeat, s->htarg, s->hhash, s->hcount, s->featb, s->out, s->expl, s->elig, s->stats, s->inv};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete s;
}

HarvestSink surrogate_harvest_sink(marl_ctx* ctx, const CellParams& p) {
  HarvestSink h{};
  SurrogateState* s = ctx->surr;
  if (!s || !s->harvesting || ctx->kin_mode != MARL_KINETICS_PLAIN) return h;
  h.feat = s->hfeat;
  h.targ = s->htarg;
  h.hash = s->hhash;
  h.count = s->hcount;
  h.cap = s->hcap;
  h.threshold = s->hthreshold;
  h.tick = p.stream;
  return h;
}

int surrogate_level(marl_ctx* ctx, const CellParams& p, const int32_t* order, int cnt, marl_cell_state* cells,
                    const marl_ruleset* rules, float* field) {
  SurrogateState* s = ctx->surr;
  if (!s->trained) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "surrogate mode needs a trained surrogate");
  if (int rc = ensure_level_scratch(ctx, cnt)) return rc;
  hipLaunchKernelGGL(k_surr_features, dim3(gs(cnt, kSBlock)), dim3(kSBlock), 0, ctx->stream, order, cnt, cells,
                     rules, field, ctx->light, ctx->nbmask, p, s->mlp.norm, s->inv, s->zmax, s->featb, s->expl,
                     s->elig, s->stats);
  MARL_HIP(ctx, hipGetLastError());
  if (int rc = mlp_infer_bf16(ctx, s->mlp, cnt, s->featb, s->out)) return rc;
  hipLaunchKernelGGL(k_surr_apply, dim3(gs(cnt, kSBlock)), dim3(kSBlock), 0, ctx->stream, order, cnt, cells, rules,
                     field, ctx->light, ctx->nbmask, ctx->events, p, s->out, s->expl, s->elig, s->mlp.norm,
                     ctx->kin_mode,
                     s->margin, s->dead_zone, s->shadow_threshold, s->stats);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int marl_set_kinetics_mode(marl_ctx* ctx, int32_t mode) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (mode != MARL_KINETICS_PLAIN && mode != MARL_KINETICS_SURROGATE && mode != MARL_KINETICS_SHADOW)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "unknown kinetics mode");
  if (mode != MARL_KINETICS_PLAIN && (!ctx->surr || !ctx->surr->trained))
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "surrogate modes need a trained surrogate");
  ctx->kin_mode = mode;
  return MARL_OK;
}

int marl_harvest_begin(marl_ctx* ctx, int64_t capacity, float sample_rate) {
  if (!ctx || capacity <= 0 || !(sample_rate > 0.0f) || sample_rate > 1.0f) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "bad harvest arguments");
  if (int rc = ensure_surr(ctx)) return rc;
  SurrogateState* s = ctx->surr;
  if (capacity > s->hcap) {
    if (s->hfeat) (void)hipFree(s->hfeat);
    if (s->htarg) (void)hipFree(s->htarg);
    if (s->hhash) (void)hipFree(s->hhash);
    s->hfeat = nullptr;
    s->htarg = nullptr;
    s->hhash = nullptr;
    MARL_HIP(ctx, hipMalloc(&s->hfeat, sizeof(float) * capacity * kFeat));
    MARL_HIP(ctx, hipMalloc(&s->htarg, sizeof(float) * capacity * kTarg));
    MARL_HIP(ctx, hipMalloc(&s->hhash, sizeof(uint64_t) * capacity));
    s->hcap = capacity;
  }
  MARL_HIP(ctx, hipMemsetAsync(s->hcount, 0, sizeof(unsigned long long), ctx->stream));
  s->hthreshold = (uint64_t)((double)sample_rate * 4294967296.0);
  s->harvesting = true;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_harvest_end(marl_ctx* ctx, int64_t* n_pairs) {
  if (!ctx || !ctx->surr) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no harvest in progress");
  SurrogateState* s = ctx->surr;
  s->harvesting = false;
  unsigned long long c = 0;
  MARL_HIP(ctx, hipMemcpyAsync(&c, s->hcount, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (n_pairs) *n_pairs = (int64_t)std::min<unsigned long long>(c, (unsigned long long)s->hcap);
  return MARL_OK;
}

static int64_t harvested(marl_ctx* ctx) {
  unsigned long long c = 0;
  if (hipMemcpyAsync(&c, ctx->surr->hcount, sizeof(c), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return -1;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
  return (int64_t)std::min<unsigned long long>(c, (unsigned long long)ctx->surr->hcap);
}

int marl_harvest_download(marl_ctx* ctx, float* feat, float* targ, uint64_t* hash, size_t cap, int64_t* n_out) {
  if (!ctx || !ctx->surr) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no harvested pairs");
  const int64_t n = harvested(ctx);
  if (n < 0) return marl_fail(ctx, MARL_ERR_TRANSFER, "reading harvest count failed");
  if (n_out) *n_out = n;
  const size_t m = std::min<size_t>(cap, (size_t)n);
  if (m == 0) return MARL_OK;
  SurrogateState* s = ctx->surr;
  if (feat) MARL_HIP(ctx, xfer(ctx, feat, s->hfeat, sizeof(float) * m * kFeat, hipMemcpyDeviceToHost));
  if (targ) MARL_HIP(ctx, xfer(ctx, targ, s->htarg, sizeof(float) * m * kTarg, hipMemcpyDeviceToHost));
  if (hash) MARL_HIP(ctx, xfer(ctx, hash, s->hhash, sizeof(uint64_t) * m, hipMemcpyDeviceToHost));
  return MARL_OK;
}

int marl_harvest_upload(marl_ctx* ctx, const float* feat, const float* targ, const uint64_t* hash, size_t n) {
  if (!ctx || !feat || !targ || !hash || n == 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "bad pairs");
  if (int rc = marl_harvest_begin(ctx, (int64_t)n, 1.0f)) return rc;
  SurrogateState* s = ctx->surr;
  s->harvesting = false;
  MARL_HIP(ctx, xfer(ctx, s->hfeat, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice));
  MARL_HIP(ctx, xfer(ctx, s->htarg, targ, sizeof(float) * n * kTarg, hipMemcpyHostToDevice));
  MARL_HIP(ctx, xfer(ctx, s->hhash, hash, sizeof(uint64_t) * n, hipMemcpyHostToDevice));
  const unsigned long long c = n;
  MARL_HIP(ctx, xfer(ctx, s->hcount, &c, sizeof(c), hipMemcpyHostToDevice));
  return MARL_OK;
}

int marl_surrogate_param_count(void) { return (int)mlp_layout().total; }

int marl_surrogate_train(marl_ctx* ctx, const marl_train_config* cfg, marl_train_report* rep) {
  if (!ctx || !cfg || !ctx->surr) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "train needs harvested pairs");
  if (cfg->steps < 0 || cfg->batch <= 0 || cfg->batch > (1 << 20))
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "bad training configuration");
  const auto t0 = std::chrono::steady_clock::now();
  SurrogateState* s = ctx->surr;
  int64_t n = harvested(ctx);
  if (n <= 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no harvested pairs");
  if (cfg->max_pairs > 0) n = std::min(n, cfg->max_pairs);
  // training / validation indices: every 20th pair validates
  std::vector<int32_t> tr, va;
  for (int64_t j = 0; j < n; ++j) (j % 20 == 19 ? va : tr).push_back((int32_t)j);
  if (tr.empty()) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no training pairs");
  if (va.empty()) va.push_back(tr.back());
  int32_t *dtr = nullptr, *dva = nullptr;
  double* dacc = nullptr;
  struct Free {
    std::vector<void*> p;
    ~Free() {
      for (void* q : p)
        if (q) (void)hipFree(q);
    }
  } fr;
  MARL_HIP(ctx, hipMalloc(&dtr, sizeof(int32_t) * tr.size()));
  fr.p.push_back(dtr);
  MARL_HIP(ctx, hipMalloc(&dva, sizeof(int32_t) * va.size()));
  fr.p.push_back(dva);
  MARL_HIP(ctx, hipMalloc(&dacc, sizeof(double) * 2 * (kFeat + kTarg)));
  fr.p.push_back(dacc);
  MARL_HIP(ctx, xfer(ctx, dtr, tr.data(), sizeof(int32_t) * tr.size(), hipMemcpyHostToDevice));
  MARL_HIP(ctx, xfer(ctx, dva, va.data(), sizeof(int32_t) * va.size(), hipMemcpyHostToDevice));
  const int B = cfg->batch;
  if (int rc = mlp_alloc(ctx, s->mlp, B)) return rc;
  // normalisation from the training pairs
  MARL_HIP(ctx, hipMemsetAsync(dacc, 0, sizeof(double) * 2 * (kFeat + kTarg), ctx->stream));
  hipLaunchKernelGGL(k_colstats, dim3(std::min<int64_t>(gs(tr.size()), 1024)), dim3(256), 0, ctx->stream, dtr,
                     (int)tr.size(), s->hfeat, s->htarg, dacc);
  MARL_HIP(ctx, hipGetLastError());
  std::vector<double> acc(2 * (kFeat + kTarg));
  MARL_HIP(ctx, xfer(ctx, acc.data(), dacc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost));
  std::vector<float> norm(2 * (kFeat + kTarg));
  const double nt = (double)tr.size();
  for (int k = 0; k < kFeat + kTarg; ++k) {
    const double mean = acc[k] / nt;
    const double var = std::max(acc[kFeat + kTarg + k] / nt - mean * mean, 0.0);
    double sd = std::sqrt(var);
    const bool bin = k < kFeat && binary_feature(k);
    if (bin) sd = 1.0;
    if (k < kFeat) {
      if (sd < 1e-6) sd = 1.0;
    } else if (sd < 1e-12 || k - kFeat == 15) {
      // constant targets, and internal[15] (set to the voxel's light by the
      // tick itself), are not learned: std 0 makes the output their mean
      sd = 0.0;
    }
    const int base = k < kFeat ? 0 : 2 * kFeat;
    const int kk = k < kFeat ? k : k - kFeat;
    const int width = k < kFeat ? kFeat : kTarg;
    norm[base + kk] = bin || (k == kFeat + 15) ? 0.0f : (float)mean;
    norm[base + width + kk] = (float)sd;
  }
  if (int rc = upload_norm(ctx, s, norm)) return rc;
  // initialisation: He-uniform weights from the seed, zero biases (fresh model only)
  const MlpLayout L = mlp_layout();
  if (!s->trained || cfg->reinit) {
    std::vector<float> w(L.total, 0.0f);
    MarlRng rng;
    rng.init(cfg->seed, 0x1417ull, 0);
    for (int l = 0; l < 4; ++l) {
      const float lim = std::sqrt(6.0f / (float)L.in[l]);
      for (size_t q = 0; q < (size_t)L.in[l] * L.out[l]; ++q) w[L.w[l] + q] = (2.0f * rng.u01() - 1.0f) * lim;
    }
    MARL_HIP(ctx, xfer(ctx, s->mlp.w, w.data(), sizeof(float) * L.total, hipMemcpyHostToDevice));
    MARL_HIP(ctx, hipMemsetAsync(s->mlp.m, 0, sizeof(float) * L.total, ctx->stream));
    MARL_HIP(ctx, hipMemsetAsync(s->mlp.v, 0, sizeof(float) * L.total, ctx->stream));
    s->mlp.step = 0;
  }
  float first = -1.0f, last = 0.0f;
  for (int64_t st = 0; st < cfg->steps; ++st) {
    hipLaunchKernelGGL(k_gather_batch, dim3(gs(B, kGatherRows)), dim3(256), 0, ctx->stream, dtr, (int)tr.size(), cfg->seed,
                       (uint64_t)st, B, s->hfeat, s->htarg, s->mlp.norm, s->inv, s->mlp.x, s->mlp.t, 0);
    MARL_HIP(ctx, hipGetLastError());
    if (int rc = mlp_forward_train(ctx, s->mlp, B)) return rc;
    const bool report = st == 0 || st == cfg->steps - 1;
    float loss = 0.0f;
    if (int rc = mlp_loss(ctx, s->mlp, B, report ? &loss : nullptr)) return rc;
    if (report) {
      if (first < 0.0f) first = loss;
      last = loss;
    }
    if (int rc = mlp_backward(ctx, s->mlp, B)) return rc;
    const double frac = cfg->steps > 1 ? (double)st / (double)(cfg->steps - 1) : 1.0;
    const float lr = (float)(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(M_PI * frac)));
    if (int rc = mlp_adam(ctx, s->mlp, lr)) return rc;
  }
  // validation loss on up to one batch of held-out pairs (in order)
  const int VB = (int)std::min<size_t>(va.size(), (size_t)B);
  hipLaunchKernelGGL(k_gather_batch, dim3(gs(VB, kGatherRows)), dim3(256), 0, ctx->stream, dva, (int)va.size(), cfg->seed, 0ull,
                     VB, s->hfeat, s->htarg, s->mlp.norm, s->inv, s->mlp.x, s->mlp.t, 1);
  MARL_HIP(ctx, hipGetLastError());
  if (int rc = mlp_forward_train(ctx, s->mlp, VB)) return rc;
  float vloss = 0.0f;
  if (int rc = mlp_loss(ctx, s->mlp, VB, &vloss)) return rc;
  if (int rc = mlp_pack(ctx, s->mlp)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->trained = true;
  s->zmax = cfg->zmax > 0.0f ? cfg->zmax : 6.0f;
  s->margin = cfg->energy_margin >= 0.0f ? cfg->energy_margin : 0.05f;
  if (rep) {
    rep->loss_first = first;
    rep->loss_last = last;
    rep->val_loss = vloss;
    rep->pairs = n;
    rep->pairs_train = (int64_t)tr.size();
    rep->ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return MARL_OK;
}

int marl_surrogate_get_params(marl_ctx* ctx, float* params, size_t n, float* norm, size_t n_norm) {
  if (!ctx || !ctx->surr || !ctx->surr->mlp.w) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  const MlpLayout L = mlp_layout();
  if (params) {
    if (n < L.total) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "params buffer too small");
    MARL_HIP(ctx, xfer(ctx, params, ctx->surr->mlp.w, sizeof(float) * L.total, hipMemcpyDeviceToHost));
  }
  if (norm) {
    if (n_norm < (size_t)2 * (kFeat + kTarg)) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "norm buffer too small");
    MARL_HIP(ctx, xfer(ctx, norm, ctx->surr->mlp.norm, sizeof(float) * 2 * (kFeat + kTarg), hipMemcpyDeviceToHost));
  }
  return MARL_OK;
}

int marl_surrogate_set_params(marl_ctx* ctx, const float* params, size_t n, const float* norm, size_t n_norm) {
  if (!ctx || !params) return MARL_ERR_INVALID_INPUT;
  const MlpLayout L = mlp_layout();
  if (n != L.total) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "parameter count mismatch");
  if (int rc = ensure_surr(ctx)) return rc;
  SurrogateState* s = ctx->surr;
  if (int rc = mlp_alloc(ctx, s->mlp, std::max(s->mlp.batch, 64))) return rc;
  MARL_HIP(ctx, xfer(ctx, s->mlp.w, params, sizeof(float) * L.total, hipMemcpyHostToDevice));
  MARL_HIP(ctx, hipMemsetAsync(s->mlp.m, 0, sizeof(float) * L.total, ctx->stream));
  MARL_HIP(ctx, hipMemsetAsync(s->mlp.v, 0, sizeof(float) * L.total, ctx->stream));
  s->mlp.step = 0;
  std::vector<float> nm(2 * (kFeat + kTarg));
  for (int k = 0; k < kFeat; ++k) { nm[k] = 0.0f; nm[kFeat + k] = 1.0f; }
  for (int k = 0; k < kTarg; ++k) { nm[2 * kFeat + k] = 0.0f; nm[2 * kFeat + kTarg + k] = 1.0f; }
  if (norm) {
    if (n_norm != nm.size()) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "norm size mismatch");
    std::copy(norm, norm + n_norm, nm.begin());
  }
  if (int rc = upload_norm(ctx, s, nm)) return rc;
  if (int rc = mlp_pack(ctx, s->mlp)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->trained = true;
  return MARL_OK;
}

int marl_surrogate_forward(marl_ctx* ctx, const float* x, size_t rows, float* y, int32_t fused) {
  if (!ctx || !x || !y || rows == 0) return MARL_ERR_INVALID_INPUT;
  if (!ctx->surr || !ctx->surr->trained) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  MlpState& m = ctx->surr->mlp;
  if (rows > (size_t)(1 << 20)) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "too many rows");
  if (int rc = mlp_alloc(ctx, m, (int)rows)) return rc;
  MARL_HIP(ctx, xfer(ctx, m.x, x, sizeof(float) * rows * kIn, hipMemcpyHostToDevice));
  float* dst = m.y;
  if (fused) {
    if (int rc = mlp_infer(ctx, m, (int)rows, m.x, dst)) return rc;
  } else {
    if (int rc = mlp_forward_train(ctx, m, (int)rows)) return rc;
  }
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  MARL_HIP(ctx, xfer(ctx, y, dst, sizeof(float) * rows * kOut, hipMemcpyDeviceToHost));
  return MARL_OK;
}

int marl_surrogate_grad(marl_ctx* ctx, const float* x, const float* t, size_t rows, float* loss, float* grads) {
  if (!ctx || !x || !t || rows == 0) return MARL_ERR_INVALID_INPUT;
  if (rows > (size_t)(1 << 20)) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "too many rows");
  if (!ctx->surr || !ctx->surr->trained) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  MlpState& m = ctx->surr->mlp;
  if (int rc = mlp_alloc(ctx, m, (int)rows)) return rc;
  MARL_HIP(ctx, xfer(ctx, m.x, x, sizeof(float) * rows * kIn, hipMemcpyHostToDevice));
  MARL_HIP(ctx, xfer(ctx, m.t, t, sizeof(float) * rows * kOut, hipMemcpyHostToDevice));
  if (int rc = mlp_forward_train(ctx, m, (int)rows)) return rc;
  float l = 0.0f;
  if (int rc = mlp_loss(ctx, m, (int)rows, &l)) return rc;
  if (int rc = mlp_backward(ctx, m, (int)rows)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l;
  if (grads) MARL_HIP(ctx, xfer(ctx, grads, m.g, sizeof(float) * mlp_layout().total, hipMemcpyDeviceToHost));
  return MARL_OK;
}

int marl_surrogate_adam_step(marl_ctx* ctx, const float* grads, float lr) {
  if (!ctx || !grads) return MARL_ERR_INVALID_INPUT;
  if (!ctx->surr || !ctx->surr->trained) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  MlpState& m = ctx->surr->mlp;
  MARL_HIP(ctx, xfer(ctx, m.g, grads, sizeof(float) * mlp_layout().total, hipMemcpyHostToDevice));
  if (int rc = mlp_adam(ctx, m, lr)) return rc;
  if (int rc = mlp_pack(ctx, m)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_surrogate_configure(marl_ctx* ctx, float zmax, float energy_margin, float shadow_rate) {
  if (!ctx || !ctx->surr) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  if (!(zmax > 0.0f) || energy_margin < 0.0f || shadow_rate < 0.0f || shadow_rate > 1.0f)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "bad surrogate settings");
  ctx->surr->zmax = zmax;
  ctx->surr->margin = energy_margin;
  ctx->surr->shadow_threshold = (uint64_t)((double)shadow_rate * 4294967296.0);
  return MARL_OK;
}

int marl_surrogate_time_infer(marl_ctx* ctx, int64_t rows, int32_t reps, float* ms_out) {
  if (!ctx || rows <= 0 || rows > (1 << 24) || reps <= 0 || !ms_out) return MARL_ERR_INVALID_INPUT;
  if (!ctx->surr || !ctx->surr->trained) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no surrogate");
  SurrogateState* s = ctx->surr;
  if (int rc = ensure_level_scratch(ctx, rows)) return rc;
  hipLaunchKernelGGL(k_fill_timing_input, dim3(gs((size_t)rows * kFeat)), dim3(256), 0, ctx->stream,
                     (size_t)rows * kFeat, s->featb);
  MARL_HIP(ctx, hipGetLastError());
  if (int rc = mlp_infer_bf16(ctx, s->mlp, (int)rows, s->featb, s->out)) return rc;  // warm-up
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int r = 0; r < reps; ++r)
    if (int rc = mlp_infer_bf16(ctx, s->mlp, (int)rows, s->featb, s->out)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.0f;
  MARL_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *ms_out = ms / (float)reps;
  return MARL_OK;
}

int marl_surrogate_set_dead_zone(marl_ctx* ctx, float dead_zone) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (!(dead_zone >= 0.0f) || !std::isfinite(dead_zone))
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "dead zone must be finite and >= 0");
  if (int rc = ensure_surr(ctx)) return rc;
  ctx->surr->dead_zone = dead_zone;
  return MARL_OK;
}

int marl_surrogate_stats(marl_ctx* ctx, marl_surr_stats* out) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  memset(out, 0, sizeof(*out));
  if (!ctx->surr) return MARL_OK;
  unsigned long long st[ST_COUNT];
  MARL_HIP(ctx, hipMemcpyAsync(st, ctx->surr->stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  out->cells = (int64_t)st[ST_CELLS];
  out->surrogate = (int64_t)st[ST_SURR];
  out->fallback_topology = (int64_t)st[ST_FB_TOPO];
  out->fallback_ood = (int64_t)st[ST_FB_OOD];
  out->fallback_margin = (int64_t)st[ST_FB_MARGIN];
  out->shadow = (int64_t)st[ST_SHADOW];
  out->shadow_within_1e3 = (int64_t)st[ST_WITHIN];
  out->shadow_fate_agree = (int64_t)st[ST_FATE_AGREE];
  const uint32_t mb = (uint32_t)st[ST_MAXERR];
  memcpy(&out->shadow_max_rel_err, &mb, 4);
  for (int b = 0; b < kErrBins; ++b) out->err_hist[b] = (int64_t)st[ST_HIST + b];
  for (int k = 0; k < 16; ++k) out->shadow_within_1e3_species[k] = (int64_t)st[ST_SPECIES + k];
  for (int e = 0; e < 4; ++e) {
    out->fate_exact[e] = (int64_t)st[ST_FATE_EXACT + e];
    out->fate_surrogate[e] = (int64_t)st[ST_FATE_SURR + e];
  }
  return MARL_OK;
}

int marl_surrogate_reset_stats(marl_ctx* ctx) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (!ctx->surr) return MARL_OK;
  MARL_HIP(ctx, hipMemsetAsync(ctx->surr->stats, 0, sizeof(unsigned long long) * ST_COUNT, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}


}  // extern "C"
