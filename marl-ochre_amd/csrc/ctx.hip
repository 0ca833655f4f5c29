// ctx.hip — context, buffers, transfers and the field entry points of the C ABI.
//
// Reference seam: crates/marl-gpu/src/field_diffusion.rs:31-268 (GpuFieldDiffuser)
// and crates/marl-gpu/src/context.rs:4-54 (GpuError, GpuContext).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "marl_internal.h"

int marl_fail(marl_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int marl_hip_fail(marl_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return marl_fail(ctx, MARL_ERR_TRANSFER, m);
}

StencilParams marl_stencil_params(const marl_sim_params* sim, int apply_sources) {
  StencilParams p;
  p.dt_sub = sim->dt / (float)sim->diffusion_substeps;  // field.rs:213
  p.alpha_eps = sim->alpha_eps;
  p.k_eps = sim->k_eps;
  for (int s = 0; s < MARL_S_EXT; ++s) {
    p.d_voxel[s] = sim->d_voxel[s];
    p.lambda_decay[s] = sim->lambda_decay[s];
  }
  p.apply_sources = apply_sources;
  p.src_ox = sim->source_rate_oxidant;
  p.src_c = sim->source_rate_carbon;
  p.src_red = sim->source_rate_reductant;
  p.c_max = sim->c_max;
  p.fuse_light = 0;
  p.l_i0 = p.l_e_cell = p.l_nchem = p.l_floor = 0.0f;
  p.light = nullptr;
  p.view = nullptr;
  return p;
}

int marl_check_params(marl_ctx* ctx, const marl_sim_params* sim) {
  if (!sim) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "invalid GPU diffuser input: null sim params");
  if (sim->diffusion_substeps < 0)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "invalid GPU diffuser input: negative diffusion_substeps");
  return MARL_OK;
}

namespace {
// Why the calling thread's last marl_create failed (there is no context to
// hold the message).
thread_local std::string t_create_err;
}  // namespace

// Fault injection for tests of a caller's fallback path (the reference's
// engine falls back to CPU diffusion when the GPU seam fails,
// crates/marl-sim/src/lib.rs:133-146, 180-191).  MARL_HIP_FAULT=init makes
// marl_create fail with DEVICE_INIT before it touches a device;
// MARL_HIP_FAULT=tick:<n> makes the n-th marl_diffuse_tick or marl_tick call
// of each context fail with TRANSFER before it changes any state, so the
// strong guarantee (host buffers and resident state untouched) can be tested.
// The variable is read at every call.
bool marl_fault_init() {
  const char* e = getenv("MARL_HIP_FAULT");
  return e && strcmp(e, "init") == 0;
}

int marl_fault_tick(marl_ctx* ctx) {
  const char* e = getenv("MARL_HIP_FAULT");
  if (!e || strncmp(e, "tick:", 5) != 0) return MARL_OK;
  const long n = strtol(e + 5, nullptr, 10);
  if (n > 0 && ++ctx->fault_calls == n)
    return marl_fail(ctx, MARL_ERR_TRANSFER, "injected fault (MARL_HIP_FAULT=" + std::string(e) + ")");
  return MARL_OK;
}

extern "C" {

const char* marl_version(void) { return MARL_VERSION; }

int marl_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}


const char* marl_last_error(const marl_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return t_create_err.empty() ? "null context" : t_create_err.c_str();
}

int marl_create(const marl_dims* dims, marl_ctx** out) {
  t_create_err.clear();
  if (!out) {
    t_create_err = "marl_create: null output pointer";
    return MARL_ERR_INVALID_INPUT;
  }
  *out = nullptr;
  if (!dims || dims->x <= 0 || dims->y <= 0 || dims->z <= 0 || dims->x % 32 != 0 || dims->x > 65535 ||
      dims->y > 65535 || dims->z > 65535) {
    t_create_err = "marl_create: grid must be positive, at most 65535 per axis, with x a multiple of 32";
    return MARL_ERR_INVALID_INPUT;
  }
  // kernels address one field buffer with 32-bit byte offsets
  if ((uint64_t)dims->x * dims->y * dims->z * MARL_S_EXT * 4ull > 0x7fffffffull) {
    t_create_err = "marl_create: one field buffer would exceed 2 GiB";
    return MARL_ERR_INVALID_INPUT;
  }
  if (marl_fault_init()) {
    t_create_err = "marl_create: injected fault (MARL_HIP_FAULT=init)";
    return MARL_ERR_DEVICE_INIT;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    t_create_err = "marl_create: no GPU visible";
    return MARL_ERR_NO_DEVICE;
  }
  marl_ctx* c = new marl_ctx();
  c->d = *dims;
  c->nvox = (size_t)dims->x * dims->y * dims->z;
  c->nfield = c->nvox * MARL_S_EXT;
  c->xw = dims->x / 32;
  c->occ_words = c->nvox / 32;
  auto fail = [&](hipError_t e, const char* what) {
    t_create_err = std::string("marl_create: ") + what + ": " + hipGetErrorString(e);
    marl_destroy(c);
    return MARL_ERR_DEVICE_INIT;
  };
  hipError_t e;
  if ((e = hipSetDevice(0)) != hipSuccess) return fail(e, "hipSetDevice");
  if (marl_noalloc_init_device(0) != MARL_OK) return fail(hipErrorInvalidSymbol, "scratch symbol address");
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
  for (int i = 0; i < 2; ++i)
    if ((e = hipMalloc(&c->field[i], c->nfield * sizeof(float))) != hipSuccess) return fail(e, "field");
  if ((e = hipMalloc(&c->occ, c->occ_words * sizeof(uint32_t))) != hipSuccess) return fail(e, "occ");
  if ((e = hipMalloc(&c->light, c->nvox * sizeof(float))) != hipSuccess) return fail(e, "light");
  if ((e = hipMemset(c->field[0], 0, c->nfield * sizeof(float))) != hipSuccess) return fail(e, "memset");
  if ((e = hipMemset(c->field[1], 0, c->nfield * sizeof(float))) != hipSuccess) return fail(e, "memset");
  if ((e = hipMemset(c->occ, 0, c->occ_words * sizeof(uint32_t))) != hipSuccess) return fail(e, "memset");
  if ((e = hipMemset(c->light, 0, c->nvox * sizeof(float))) != hipSuccess) return fail(e, "memset");
  if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return fail(e, "event");
  if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return fail(e, "event");
  if ((e = hipHostMalloc((void**)&c->pinned, 4096, hipHostMallocDefault)) != hipSuccess) return fail(e, "pinned");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "sync");
  c->tile = marl_default_tile(c->d);
  *out = c;
  return MARL_OK;
}

void marl_destroy(marl_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->aux) (void)hipStreamSynchronize(c->aux);
  surrogate_free(c->surr);
  for (int k = 0; k < 2; ++k) {
    if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
    if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
  }
  if (c->tdev) (void)hipFree(c->tdev);
  void* bufs[] = {c->field[0], c->field[1], c->occ, c->light, c->staging, c->occ_bytes, c->cells[0], c->cells[1],
                  c->rules[0], c->rules[1], c->cell_at, c->nbmask, c->level, c->order, c->events, c->flag_scan,
                  c->flag_scan2, c->misc, c->level_count, c->claim, c->birth_cells, c->birth_rules,
                  c->birth_flag, c->scan_tmp, c->budget, c->preds, c->npred, c->phase_budget_dev};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->tick_ring) (void)hipHostFree(c->tick_ring);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (auto x : c->xev)
    if (x) (void)hipEventDestroy(x);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int marl_sync(marl_ctx* ctx) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

}  // extern "C"

// The second stream and its fork/join events, created on first use (contexts
// that only ever tick inside an ensemble never need them).
int marl_ensure_aux(marl_ctx* ctx) {
  if (ctx->aux) return MARL_OK;
  for (auto& x : ctx->xev)
    if (!x) MARL_HIP(ctx, hipEventCreateWithFlags(&x, hipEventDisableTiming));
  MARL_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
  return MARL_OK;
}

int marl_ensure_staging(marl_ctx* ctx) {
  if (!ctx->staging) MARL_HIP(ctx, hipMalloc(&ctx->staging, ctx->nfield * sizeof(float)));
  if (!ctx->occ_bytes) MARL_HIP(ctx, hipMalloc(&ctx->occ_bytes, ctx->nvox + 64));
  return MARL_OK;
}

static int check_len(marl_ctx* ctx, size_t got, size_t want, const char* what) {
  if (got != want) {
    char buf[160];
    snprintf(buf, sizeof buf, "invalid GPU diffuser input: %s has %zu elements, expected %zu", what, got, want);
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, buf);
  }
  return MARL_OK;
}

// Device->host copy with the strong guarantee: the destination is written only
// after the whole transfer succeeded.
static int d2h_checked(marl_ctx* ctx, void* host, const void* dev, size_t bytes) {
  void* tmp = nullptr;
  hipError_t e = hipHostMalloc(&tmp, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return marl_hip_fail(ctx, e, "hipHostMalloc(readback)");
  e = hipMemcpyAsync(tmp, dev, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipHostFree(tmp);
    return marl_hip_fail(ctx, e, "readback");
  }
  memcpy(host, tmp, bytes);
  (void)hipHostFree(tmp);
  return MARL_OK;
}

extern "C" {

int marl_upload_field_aos(marl_ctx* ctx, const float* aos, size_t n) {
  if (!ctx || !aos) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nfield, "field")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->staging, aos, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_download_field_aos(marl_ctx* ctx, float* aos, size_t n) {
  if (!ctx || !aos) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nfield, "field")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  if (int rc = marl_transpose(ctx, false, ctx->field[ctx->cur])) return rc;
  return d2h_checked(ctx, aos, ctx->staging, n * sizeof(float));
}

int marl_upload_field_soa(marl_ctx* ctx, const float* soa, size_t n) {
  if (!ctx || !soa) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nfield, "field")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->staging, soa, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if (int rc = marl_planes(ctx, true, ctx->field[ctx->cur])) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_download_field_soa(marl_ctx* ctx, float* soa, size_t n) {
  if (!ctx || !soa) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nfield, "field")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  if (int rc = marl_planes(ctx, false, ctx->field[ctx->cur])) return rc;
  return d2h_checked(ctx, soa, ctx->staging, n * sizeof(float));
}

int marl_upload_occupancy(marl_ctx* ctx, const uint8_t* occ, size_t n) {
  if (!ctx || !occ) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nvox, "occupancy")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->occ_bytes, occ, n, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = marl_pack_occ(ctx)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_download_occupancy(marl_ctx* ctx, uint8_t* occ, size_t n) {
  if (!ctx || !occ) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nvox, "occupancy")) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  if (int rc = marl_unpack_occ(ctx)) return rc;
  return d2h_checked(ctx, occ, ctx->occ_bytes, n);
}

int marl_download_light(marl_ctx* ctx, float* light, size_t n) {
  if (!ctx || !light) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null argument");
  if (int rc = check_len(ctx, n, ctx->nvox, "light")) return rc;
  return d2h_checked(ctx, light, ctx->light, n * sizeof(float));
}

int marl_apply_boundary_sources(marl_ctx* ctx, const marl_sim_params* sim) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (int rc = marl_launch_boundary(ctx, sim)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_diffuse_substeps(marl_ctx* ctx, const marl_sim_params* sim, int32_t n) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (sim->diffusion_substeps == 0 || n <= 0) return MARL_OK;
  if (int rc = marl_launch_substeps(ctx, sim, n, false, ctx->tile)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

// Drop-in: ≙ GpuFieldDiffuser::diffuse_tick_with_cells (field_diffusion.rs:158-268).
int marl_diffuse_tick(marl_ctx* ctx, float* aos, size_t n_floats, const uint8_t* occ, size_t n_voxels,
                      const marl_sim_params* sim) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (!aos || !occ) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "invalid GPU diffuser input: null buffer");
  if (int rc = check_len(ctx, n_floats, ctx->nfield, "field")) return rc;      // :164-169
  if (int rc = check_len(ctx, n_voxels, ctx->nvox, "occupancy")) return rc;    // :170-175
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (sim->diffusion_substeps == 0) return MARL_OK;                             // :176-178
  if (int rc = marl_fault_tick(ctx)) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->staging, aos, n_floats * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  MARL_HIP(ctx, hipMemcpyAsync(ctx->occ_bytes, occ, n_voxels, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) return rc;
  if (int rc = marl_pack_occ(ctx)) return rc;
  if (int rc = marl_launch_substeps(ctx, sim, sim->diffusion_substeps, false, ctx->tile)) return rc;
  if (int rc = marl_transpose(ctx, false, ctx->field[ctx->cur])) return rc;
  return d2h_checked(ctx, aos, ctx->staging, n_floats * sizeof(float));         // :270-292
}

int marl_light_update(marl_ctx* ctx, const marl_sim_params* sim) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (int rc = marl_launch_light(ctx, sim)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_time_diffuse(marl_ctx* ctx, const marl_sim_params* sim, int32_t n, float* ms) {
  if (!ctx || !ms) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (sim->diffusion_substeps == 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "diffusion_substeps == 0");
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (int rc = marl_launch_substeps(ctx, sim, n, false, ctx->tile)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return MARL_OK;
}

int marl_time_field_phase(marl_ctx* ctx, const marl_sim_params* sim, int32_t n_ticks, int32_t with_sources,
                          float* ms) {
  if (!ctx || !ms || n_ticks < 0) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < n_ticks; ++k)
    if (int rc = marl_launch_field_phase(ctx, sim, with_sources != 0)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return MARL_OK;
}

int marl_time_light(marl_ctx* ctx, const marl_sim_params* sim, int32_t n, float* ms) {
  if (!ctx || !ms) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < n; ++k)
    if (int rc = marl_launch_light(ctx, sim)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return MARL_OK;
}

int marl_copy_probe(marl_ctx* ctx, int32_t n, float* ms, double* bytes) {
  if (!ctx || !ms || n <= 0) return MARL_ERR_INVALID_INPUT;
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < n; ++k)
    if (int rc = marl_launch_copy_probe(ctx)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  if (bytes) *bytes = 2.0 * (double)(ctx->nfield / 4) * 16.0;
  return MARL_OK;
}

// Sectors of a 1 GiB buffer (four times the last-level cache), allocated for
// the measurement and freed after it; contents are irrelevant.
int marl_gather_probe(marl_ctx* ctx, int32_t n, float* ms, double* bytes) {
  if (!ctx || !ms || n <= 0) return MARL_ERR_INVALID_INPUT;
  constexpr size_t kSrcBytes = (size_t)1 << 30, kThreads = (size_t)1 << 24;
  void* src = nullptr;
  float* out = nullptr;
  MARL_HIP(ctx, hipMalloc(&src, kSrcBytes));
  if (hipMalloc(&out, kThreads * sizeof(float)) != hipSuccess) {
    (void)hipFree(src);
    return marl_fail(ctx, MARL_ERR_TRANSFER, "gather probe: allocation failed");
  }
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{src, out};
  const uint32_t nsec = (uint32_t)(kSrcBytes / 64);
  double per = 0.0;
  if (int rc = marl_launch_gather_probe(ctx, src, nsec, out, kThreads, 0x9e3779b9u, &per)) return rc;  // warm-up
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < n; ++k)
    if (int rc = marl_launch_gather_probe(ctx, src, nsec, out, kThreads, (uint32_t)k * 0x85ebca6bu + 1u, &per))
      return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  if (bytes) *bytes = per;
  return MARL_OK;
}

// Dense BF16 MFMA rate: register-only chains of 32x32x16 products on every
// CU; the output buffer (one float per thread) is allocated for the
// measurement and freed after it.
int marl_mfma_probe(marl_ctx* ctx, int32_t n, float* ms, double* flop) {
  if (!ctx || !ms || n <= 0) return MARL_ERR_INVALID_INPUT;
  int cus = 0;
  MARL_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if (cus <= 0) return marl_fail(ctx, MARL_ERR_TRANSFER, "mfma probe: no compute units reported");
  float* out = nullptr;
  MARL_HIP(ctx, hipMalloc(&out, marl_mfma_probe_floats(cus) * sizeof(float)));
  struct Free {
    float* p;
    ~Free() { (void)hipFree(p); }
  } fr{out};
  double per = 0.0;
  if (int rc = marl_launch_mfma_probe(ctx, cus, out, &per)) return rc;  // warm-up
  MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < n; ++k)
    if (int rc = marl_launch_mfma_probe(ctx, cus, out, &per)) return rc;
  MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MARL_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  if (flop) *flop = per;
  return MARL_OK;
}

int marl_stencil_tile_count(void) { return kNumTiles; }

int marl_stencil_tile_substeps(int32_t tile) {
  if (tile < 0 || tile >= kNumTiles) return -MARL_ERR_INVALID_INPUT;
  return kTiles[tile].tb;
}

int marl_stencil_tile_shape(int32_t tile, int32_t* ry, int32_t* zc, int32_t* pd) {
  if (tile < 0 || tile >= kNumTiles) return MARL_ERR_INVALID_INPUT;
  if (ry) *ry = kTiles[tile].ry;
  if (zc) *zc = kTiles[tile].zc;
  if (pd) *pd = kTiles[tile].pd;
  return MARL_OK;
}

int marl_stencil_tile_split(int32_t tile, int32_t* sg, int32_t* nw) {
  if (tile < 0 || tile >= kNumTiles) return MARL_ERR_INVALID_INPUT;
  if (sg) *sg = kTiles[tile].sg;
  if (nw) *nw = kTiles[tile].nw;
  return MARL_OK;
}

int marl_set_stencil_tile(marl_ctx* ctx, int32_t tile) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (tile < 0 || tile >= kNumTiles) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "tile id out of range");
  ctx->tile = tile;
  return MARL_OK;
}

int marl_get_stencil_tile(marl_ctx* ctx, int32_t* tile) {
  if (!ctx || !tile) return MARL_ERR_INVALID_INPUT;
  *tile = ctx->tile;
  return MARL_OK;
}

// Autotuner: times every tile shape on the resident field (2 warm-up + 8 timed
// substeps each), keeps the fastest, and restores the field it started from.
int marl_autotune_stencil(marl_ctx* ctx, const marl_sim_params* sim, int32_t* tile_out) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (sim->diffusion_substeps == 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "diffusion_substeps == 0");
  if (int rc = marl_ensure_staging(ctx)) return rc;
  const int cur0 = ctx->cur;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->staging, ctx->field[cur0], ctx->nfield * sizeof(float),
                               hipMemcpyDeviceToDevice, ctx->stream));
  int best = ctx->tile;
  float best_ms = 1e30f;
  for (int t = 0; t < kNumTiles; ++t) {
    if (int rc = marl_launch_substeps(ctx, sim, 2, false, t)) return rc;
    MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (int rc = marl_launch_substeps(ctx, sim, 8, false, t)) return rc;
    MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    MARL_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    MARL_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  ctx->cur = cur0;
  MARL_HIP(ctx, hipMemcpyAsync(ctx->field[cur0], ctx->staging, ctx->nfield * sizeof(float),
                               hipMemcpyDeviceToDevice, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->tile = best;
  if (tile_out) *tile_out = best;
  return MARL_OK;
}

}  // extern "C"
