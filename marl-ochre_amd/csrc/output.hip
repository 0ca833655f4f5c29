// output.hip — the reductions and packing behind the run's outputs, on the
// resident state, so that only small results cross PCIe.
//
// Reference:
//   per-tick summary     crates/marl-output/src/data.rs:195-275 (DataLogger::log_tick),
//                        crates/marl-sim/src/stats.rs:9-82 (print_stats)
//   chemistry profile    crates/marl-output/src/data.rs:292-335 (snapshot_chemistry)
//   viewer cell records  crates/marl-output/src/binary_dump.rs:34-41, 115-120 (write_cell_dump),
//                        crates/marl-format/src/lib.rs:350-364 (ViewerCellRecord, 25 bytes packed)
//   point samples        crates/marl-sim/src/stats.rs:51-56, 86-103 (centre-column reads)
#include <algorithm>
#include <mutex>
#include <vector>

#include "marl_internal.h"

namespace {

constexpr int kSumBlocks = 256;
constexpr int kZLocal = 2048;  // z-histogram aggregated in LDS up to this depth

// FP64 partial sums of internal[0], [5], [6] per block; active reactions,
// z-thirds and per-layer counts as integer atomics (LDS-aggregated).
__global__ void __launch_bounds__(256) k_cell_summary(const marl_cell_state* __restrict__ cells,
                                                      const marl_ruleset* __restrict__ rules, int n, int Z,
                                                      double* __restrict__ partial,
                                                      unsigned long long* __restrict__ counts /* 4 */,
                                                      unsigned long long* __restrict__ zc /* Z */) {
  __shared__ double red[3][256];
  __shared__ unsigned int zl[kZLocal];
  __shared__ unsigned int cl[4];
  const bool local_z = Z <= kZLocal;
  if (local_z)
    for (int k = threadIdx.x; k < Z; k += blockDim.x) zl[k] = 0u;
  if (threadIdx.x < 4) cl[threadIdx.x] = 0u;
  __syncthreads();
  const int z_third = Z / 3;  // stats.rs:39 (as u16)
  double se = 0.0, sa = 0.0, sb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const marl_cell_state& c = cells[i];
    se += (double)c.internal[0];
    sa += (double)c.internal[5];
    sb += (double)c.internal[6];
    unsigned active = 0;
#pragma unroll
    for (int q = 0; q < MARL_R_MAX; ++q) active += fabsf(rules[i].reactions[q].v_max) > 1e-9f ? 1u : 0u;
    atomicAdd(&cl[0], active);
    const int z = c.pos[2];
    atomicAdd(&cl[z < z_third ? 1 : (z < 2 * z_third ? 2 : 3)], 1u);
    if (z < Z) {
      if (local_z) atomicAdd(&zl[z], 1u);
      else atomicAdd(&zc[z], 1ull);
    }
  }
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sa;
  red[2][threadIdx.x] = sb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
  if (threadIdx.x < 4 && cl[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)cl[threadIdx.x]);
  if (local_z)
    for (int k = threadIdx.x; k < Z; k += blockDim.x)
      if (zl[k]) atomicAdd(&zc[k], (unsigned long long)zl[k]);
}

// Mean over each XY plane of each species (and of light, column 12), FP64,
// one block per (z, column).
__global__ void __launch_bounds__(256) k_chem_profile(const float* __restrict__ field,
                                                      const float* __restrict__ light, int X, int Y, int Z,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  const int z = blockIdx.x, s = blockIdx.y;
  const size_t XY = (size_t)X * Y;
  const size_t nvox = XY * Z;
  // species s of the quad field sits every 4th float; light is dense
  const float* plane = s < MARL_S_EXT ? field + fq_index(s, (size_t)z * XY, nvox) : light + (size_t)z * XY;
  const size_t stride = s < MARL_S_EXT ? 4 : 1;
  double acc = 0.0;
  for (size_t k = threadIdx.x; k < XY; k += blockDim.x) acc += (double)plane[k * stride];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(size_t)z * (MARL_S_EXT + 1) + s] = red[0] / (double)XY;
}

// ViewerCellRecord {pos: f32[3], lineage_id: u64, starter_type: u8, energy: f32}, packed.
__global__ void k_pack_viewer(const marl_cell_state* __restrict__ cells, int n, uint8_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const marl_cell_state& c = cells[i];
  uint8_t rec[25];
  const float pos[3] = {(float)c.pos[0], (float)c.pos[1], (float)c.pos[2]};
  __builtin_memcpy(rec, pos, 12);
  __builtin_memcpy(rec + 12, &c.lineage_id, 8);
  rec[20] = c.starter_type;
  __builtin_memcpy(rec + 21, &c.internal[0], 4);
  uint8_t* o = out + (size_t)i * 25;
#pragma unroll
  for (int k = 0; k < 25; ++k) o[k] = rec[k];
}

// Point reads: q[k] = (x, y, z, species), species -1 = light.
__global__ void k_sample(const float* __restrict__ field, const float* __restrict__ light,
                         const int32_t* __restrict__ q, int n, int X, int Y, int Z, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int x = q[4 * k], y = q[4 * k + 1], z = q[4 * k + 2], s = q[4 * k + 3];
  const size_t nvox = (size_t)X * Y * Z;
  const size_t v = ((size_t)z * Y + y) * X + x;
  out[k] = s < 0 ? light[v] : field[fq_index(s, v, nvox)];
}

template <typename T>
int dev_scratch(marl_ctx* ctx, T** p, size_t count) {
  MARL_HIP(ctx, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return MARL_OK;
}

// ---- the population summary without allocation ------------------------------
// marl_cell_summary_noalloc keeps its partial sums, counters and result record
// in one file-scope device array instead of memory it allocates per call.
// Each GPU holds its own copy of the array at its own address; host code may
// use only the address hipGetSymbolAddress returns for the current device,
// looked up once per device at context creation (marl_noalloc_init_device).
struct NoAllocSummary {
  double partial[kSumBlocks * 3];                    // FP64 block partials
  unsigned long long counts[4 + kNoAllocSummaryMaxZ];  // active, thirds, per-layer
  marl_population_summary result;                    // summed on the device
  int64_t z_counts[kNoAllocSummaryMaxZ];
};
static_assert(sizeof(NoAllocSummary) <= kNoAllocScratchBytes, "summary fits the scratch array");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Counters and partials to zero (a fill kernel, never a memset node, so that
// the call stays capturable).
__global__ void k_summary_clear(NoAllocSummary* s, int nz) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 4 + nz; k += gridDim.x * blockDim.x) s->counts[k] = 0ull;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kSumBlocks * 3; k += gridDim.x * blockDim.x)
    s->partial[k] = 0.0;
}

// The partials summed in block order, one thread: the same FP64 additions in
// the same order as marl_cell_summary's host loop, so the sums are identical.
__global__ void k_summary_finish(NoAllocSummary* s, int64_t n, int nz) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    marl_population_summary r{};
    r.n_cells = n;
    for (int b = 0; b < kSumBlocks; ++b) {
      r.sum_energy += s->partial[b * 3];
      r.sum_enzyme_a += s->partial[b * 3 + 1];
      r.sum_enzyme_b += s->partial[b * 3 + 2];
    }
    r.sum_active_reactions = (int64_t)s->counts[0];
    for (int k = 0; k < 3; ++k) r.z_thirds[k] = (int64_t)s->counts[1 + k];
    s->result = r;
  }
  for (int z = blockIdx.x * blockDim.x + threadIdx.x; z < nz; z += gridDim.x * blockDim.x)
    s->z_counts[z] = (int64_t)s->counts[4 + z];
}

}  // namespace

int marl_noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return MARL_ERR_DEVICE_INIT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return MARL_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return MARL_ERR_DEVICE_INIT;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MARL_ERR_DEVICE_INIT;
  g_scratch_ptr[device] = ptr;
  return MARL_OK;
}

extern "C" {

int marl_cell_summary(marl_ctx* ctx, marl_population_summary* out, int64_t* z_counts, size_t nz) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  if (z_counts && nz < (size_t)ctx->d.z) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "z_counts shorter than Z");
  const int n = (int)ctx->n_cells;
  const int Z = ctx->d.z;
  double* partial = nullptr;
  unsigned long long* cnt = nullptr;
  if (int rc = dev_scratch(ctx, &partial, kSumBlocks * 3)) return rc;
  if (int rc = dev_scratch(ctx, &cnt, 4 + (size_t)Z)) {
    (void)hipFree(partial);
    return rc;
  }
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{partial, cnt};
  MARL_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (4 + (size_t)Z), ctx->stream));
  MARL_HIP(ctx, hipMemsetAsync(partial, 0, sizeof(double) * kSumBlocks * 3, ctx->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_summary, dim3(kSumBlocks), dim3(256), 0, ctx->stream, ctx->cells[ctx->cbuf],
                       ctx->rules[ctx->cbuf], n, Z, partial, cnt, cnt + 4);
    MARL_HIP(ctx, hipGetLastError());
  }
  std::vector<double> hp(kSumBlocks * 3);
  std::vector<unsigned long long> hc(4 + (size_t)Z);
  MARL_HIP(ctx, hipMemcpyAsync(hp.data(), partial, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipMemcpyAsync(hc.data(), cnt, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  marl_population_summary s{};
  s.n_cells = n;
  for (int b = 0; b < kSumBlocks; ++b) {  // fixed order: deterministic
    s.sum_energy += hp[b * 3];
    s.sum_enzyme_a += hp[b * 3 + 1];
    s.sum_enzyme_b += hp[b * 3 + 2];
  }
  s.sum_active_reactions = (int64_t)hc[0];
  for (int k = 0; k < 3; ++k) s.z_thirds[k] = (int64_t)hc[1 + k];
  *out = s;
  if (z_counts)
    for (int z = 0; z < Z; ++z) z_counts[z] = (int64_t)hc[4 + z];
  return MARL_OK;
}

int marl_cell_summary_noalloc(marl_ctx* ctx, marl_population_summary* out, int64_t* z_counts, size_t nz) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  const int Z = ctx->d.z;
  if (z_counts && nz < (size_t)Z) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "z_counts shorter than Z");
  if (Z > kNoAllocSummaryMaxZ)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT,
                     "marl_cell_summary_noalloc: Z above the scratch limit of " +
                         std::to_string(kNoAllocSummaryMaxZ) + " layers");
  void* base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  if (!base) return marl_fail(ctx, MARL_ERR_DEVICE_INIT, "marl_cell_summary_noalloc: scratch address not set up");
  NoAllocSummary* s = static_cast<NoAllocSummary*>(base);
  const int n = (int)ctx->n_cells;
  hipLaunchKernelGGL(k_summary_clear, dim3(8), dim3(256), 0, ctx->stream, s, Z);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_summary, dim3(kSumBlocks), dim3(256), 0, ctx->stream, ctx->cells[ctx->cbuf],
                       ctx->rules[ctx->cbuf], n, Z, s->partial, s->counts, s->counts + 4);
  hipLaunchKernelGGL(k_summary_finish, dim3(8), dim3(256), 0, ctx->stream, s, (int64_t)n, Z);
  MARL_HIP(ctx, hipGetLastError());
  MARL_HIP(ctx, hipMemcpyAsync(out, &s->result, sizeof(marl_population_summary), hipMemcpyDeviceToHost,
                               ctx->stream));
  if (z_counts)
    MARL_HIP(ctx, hipMemcpyAsync(z_counts, s->z_counts, sizeof(int64_t) * (size_t)Z, hipMemcpyDeviceToHost,
                                 ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_chemistry_profile(marl_ctx* ctx, double* out, size_t n) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  const size_t need = (size_t)ctx->d.z * (MARL_S_EXT + 1);
  if (n < need) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "profile buffer shorter than Z * 13");
  double* d = nullptr;
  if (int rc = dev_scratch(ctx, &d, need)) return rc;
  hipLaunchKernelGGL(k_chem_profile, dim3(ctx->d.z, MARL_S_EXT + 1), dim3(256), 0, ctx->stream, ctx->field[ctx->cur],
                     ctx->light, ctx->d.x, ctx->d.y, ctx->d.z, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(double) * need, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  MARL_HIP(ctx, e);
  return MARL_OK;
}

int marl_download_viewer_cells(marl_ctx* ctx, uint8_t* out, size_t cap_bytes, int64_t* n_out) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  const int64_t n = ctx->n_cells;
  if (n_out) *n_out = n;
  if (n == 0) return MARL_OK;
  if (!out || cap_bytes < (size_t)n * 25) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "buffer shorter than 25 * cells");
  uint8_t* d = nullptr;
  if (int rc = dev_scratch(ctx, &d, (size_t)n * 25)) return rc;
  hipLaunchKernelGGL(k_pack_viewer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     ctx->cells[ctx->cbuf], (int)n, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)n * 25, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  MARL_HIP(ctx, e);
  return MARL_OK;
}

int marl_sample(marl_ctx* ctx, const int32_t* queries, size_t n, float* out) {
  if (!ctx || (n && (!queries || !out))) return MARL_ERR_INVALID_INPUT;
  if (n == 0) return MARL_OK;
  for (size_t k = 0; k < n; ++k) {
    const int32_t* q = queries + 4 * k;
    if (q[0] < 0 || q[0] >= ctx->d.x || q[1] < 0 || q[1] >= ctx->d.y || q[2] < 0 || q[2] >= ctx->d.z || q[3] < -1 ||
        q[3] >= MARL_S_EXT)
      return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "sample out of range");
  }
  int32_t* dq = nullptr;
  float* dout = nullptr;
  if (int rc = dev_scratch(ctx, &dq, 4 * n)) return rc;
  if (int rc = dev_scratch(ctx, &dout, n)) {
    (void)hipFree(dq);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(dq, queries, sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->field[ctx->cur],
                       ctx->light, dq, (int)n, ctx->d.x, ctx->d.y, ctx->d.z, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dq);
  (void)hipFree(dout);
  MARL_HIP(ctx, e);
  return MARL_OK;
}

}  // extern "C"
