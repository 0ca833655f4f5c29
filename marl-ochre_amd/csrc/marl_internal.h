// marl_internal.h — shared declarations of the gfx950 MARL library (host + device).
//
// Device layout (HBM):
//   field      species quads: three planes of float4 [g][z][y][x][4], plane g holding
//              species 4g..4g+3 of every voxel (fq_index), ping-pong pair
//   occupancy  1 bit per voxel, u32 words along x: word (z*Y + y)*XW + x/32, XW = X/32
//   light      f32 [z][y][x]
//   cells      AoS marl_cell_state (96 B) + marl_ruleset (576 B) in `cells` Vec order
//   cell_at    i32 per voxel: index of the cell in the voxel, -1 when free
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>
#include <vector>

#include "../../include/marl_hip.h"

// Species s of voxel v in the resident field: float index of the quad layout.
// A voxel's four species of one quad are adjacent (one 16-byte load), and the
// quad planes are nvox float4 apart.
__host__ __device__ inline size_t fq_index(int s, size_t v, size_t nvox) {
  return ((size_t)(s >> 2) * nvox + v) * 4 + (size_t)(s & 3);
}

#define MARL_VERSION "marl-ochre-amd 0.1.0 (gfx950)"

// ---------------------------------------------------------------------------
// Math shared by every kernel that must agree bit-for-bit with the CPU oracle.
// f32 arithmetic in the reference's operator order; this library is built with
// -ffp-contract=off and IEEE division/sqrt, denormals preserved.

// exp(x) by Cody-Waite reduction (ln2 split 0.693145751953125 + 1.4286068e-6)
// and the degree-7 Taylor polynomial of e^r on |r| <= ln2/2; results below
// FLT_MIN flush to zero.  Within 2 ulp of libm expf.
__host__ __device__ inline float marl_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283f) return __builtin_huge_valf();
  if (x < -87.33654f) return 0.0f;
  float t = x * 1.44269504f;
  float n = floorf(t + 0.5f);
  float r = x - n * 0.693145751953125f;
  r = r - n * 1.428606765330187e-06f;
  float p = 1.98412698e-04f;
  p = p * r + 1.38888889e-03f;
  p = p * r + 8.33333333e-03f;
  p = p * r + 4.16666667e-02f;
  p = p * r + 1.66666667e-01f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int ni = (int)n;
  if (ni > 127) { p = p * 2.0f; ni -= 1; }
  uint32_t b = (uint32_t)(ni + 127) << 23;
  float s;
  __builtin_memcpy(&s, &b, 4);
  return p * s;
}

// log(x) for normal x > 0: x = m*2^e with m in [sqrt(1/2), sqrt(2)], then
// log m = 2 atanh(f), f = (m-1)/(m+1), odd series to f^9.
__host__ __device__ inline float marl_logf(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_huge_valf() : __builtin_nanf("");
  if (x == __builtin_huge_valf()) return x;
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  int e = (int)((b >> 23) & 0xff) - 127;
  uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  float m;
  __builtin_memcpy(&m, &mb, 4);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float f = (m - 1.0f) / (m + 1.0f);
  float f2 = f * f;
  float s = 1.11111111e-01f;
  s = s * f2 + 1.42857143e-01f;
  s = s * f2 + 2.00000000e-01f;
  s = s * f2 + 3.33333333e-01f;
  s = s * f2;
  float twof = 2.0f * f;
  float lm = twof + twof * s;
  float fe = (float)e;
  return fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
}

// ---------------------------------------------------------------------------
// Philox4x64-10 counter-based RNG (Salmon et al. SC'11), the same bijection as
// numpy.random.Philox.  Stream layout: key (seed, stream), counter
// (draw/8, entity, 0, 0); u32 draw j is word (j>>1)&3, low half when j even.

#define MARL_TAG_SEEDING 1ull
#define MARL_TAG_DIVISION 2ull
#define MARL_TAG_HGT 3ull
#define MARL_TAG_SCENARIO 3ull

__host__ __device__ inline void marl_mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  *hi = __umul64hi(a, b);
  *lo = a * b;
#else
  __uint128_t p = (__uint128_t)a * b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
#endif
}

__host__ __device__ inline void marl_philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    uint64_t hi0, lo0, hi1, lo1;
    marl_mulhilo64(0xD2E7470EE14C6C93ull, c[0], &hi0, &lo0);
    marl_mulhilo64(0xCA5A826395121157ull, c[2], &hi1, &lo1);
    uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

// Sequential reader of one entity's stream; caches the current 8-draw block.
struct MarlRng {
  uint64_t seed, stream, entity, draw;
  uint64_t blk[4];
  uint64_t blk_id;
  // optional: the stream's blocks [0, tab_n) computed ahead (4 words each,
  // e.g. by the other lanes of a wave); later blocks are computed here
  const uint64_t* tab;
  uint32_t tab_n;
  __host__ __device__ void init(uint64_t s, uint64_t st, uint64_t e, uint64_t d0 = 0) {
    seed = s; stream = st; entity = e; draw = d0; blk_id = ~0ull; tab = nullptr; tab_n = 0;
  }
  __host__ __device__ uint32_t u32() {
    uint64_t b = draw >> 3;
    if (b != blk_id) {
      if (b < tab_n) {
        for (int k = 0; k < 4; ++k) blk[k] = tab[4 * b + k];
      } else {
        blk[0] = b; blk[1] = entity; blk[2] = 0; blk[3] = 0;
        marl_philox4x64(blk, seed, stream);
      }
      blk_id = b;
    }
    uint64_t w = blk[(draw >> 1) & 3];
    uint32_t r = (draw & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
    ++draw;
    return r;
  }
  // uniform f32 in [0,1) from the top 24 bits
  __host__ __device__ float u01() { return (float)(u32() >> 8) * 0x1p-24f; }
  // integer in [0,k) by the widening multiply (k >= 1)
  __host__ __device__ uint32_t range(uint32_t k) { return (uint32_t)(((uint64_t)u32() * k) >> 32); }
  __host__ __device__ uint64_t u64() {
    uint64_t lo = u32();
    uint64_t hi = u32();
    return lo | (hi << 32);
  }
  // 0 + sd * N(0,1), Marsaglia polar method
  __host__ __device__ float normal(float sd) {
    float u, v, s;
    do {
      u = 2.0f * u01() - 1.0f;
      v = 2.0f * u01() - 1.0f;
      s = u * u + v * v;
    } while (!(s < 1.0f && s > 0.0f));
    float z = u * sqrtf((-2.0f * marl_logf(s)) / s);
    return 0.0f + sd * z;
  }
};

// ---------------------------------------------------------------------------
// Stencil parameters passed by value.
struct StencilParams {
  float dt_sub, alpha_eps, k_eps;
  float d_voxel[MARL_S_EXT];
  float lambda_decay[MARL_S_EXT];
  // boundary sources fused into the first substep of a tick (field.rs:231-250)
  int32_t apply_sources;
  float src_ox, src_c, src_red, c_max;
  // light scan fused into the last substep of a tick (light.rs:29-59); only
  // for tiles whose z-chunk spans the whole column
  int32_t fuse_light;  // last substep's light work: 0 none, 1 the scan, 2 compact species 4 into the light plane
  float l_i0, l_e_cell, l_nchem, l_floor;
  float* light;
  // last substep of a tick: the new field goes into this AoS view (48 B per
  // voxel, the cell pass's layout) instead of the quads, when not null
  float* view;
};

// ---------------------------------------------------------------------------
// XCD-aware block order: hardware dispatches block b to XCD b % 8; give each
// XCD a contiguous run of logical blocks so that blocks touching neighbouring
// data (stencil tiles sharing halo lines, spatially adjacent cells) sit in the
// same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int nb8 = nb & ~7;
  return b < nb8 ? (b & 7) * (nb8 >> 3) + (b >> 3) : b;
}

// ---------------------------------------------------------------------------
// One column of an ensemble as the batched field kernels see it.
struct EnsColumn {
  float* f[2];           // the column's field ping-pong buffers (species quads)
  const uint32_t* occ;   // its occupancy bits
  float* light;          // its light plane
  int X, Y, Z, XW;
  int ntx, nty, ntz, zc; // stencil tiles of the ensemble's tile shape
  int tile_base;         // first logical stencil tile (species groups included)
  int cur;               // buffer holding the field at the start of a tick
  int light_base;        // first 64-lane light block
  int lbx;               // light blocks along x
};

// ---------------------------------------------------------------------------
// Host context.
struct SurrogateState;
void surrogate_free(SurrogateState* s);

struct marl_ctx {
  marl_dims d{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t nvox = 0, nfield = 0, occ_words = 0;
  int xw = 0;  // occupancy words per row

  float* field[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* occ = nullptr;
  float* light = nullptr;
  float* staging = nullptr;  // AoS staging (nfield floats), allocated on first use
  // the last field phase left the new field in the AoS view (staging) only;
  // the quads field[cur] are stale until the cell pass transposes the view back
  bool view_fresh = false;
  uint8_t* occ_bytes = nullptr;
  int tile = -1;             // stencil tile id; -1 = default

  // cells
  int64_t n_cells = 0, cap = 0;
  marl_cell_state* cells[2] = {nullptr, nullptr};
  marl_ruleset* rules[2] = {nullptr, nullptr};
  int cbuf = 0;
  int32_t* cell_at = nullptr;
  // per-cell scratch (capacity cap)
  uint8_t* nbmask = nullptr;   // 6-bit empty-neighbour mask
  int32_t* level = nullptr;
  int32_t* preds = nullptr;    // conflict predecessors: 18 rows of cap entries (row q = q-th predecessor)
  uint8_t* npred = nullptr;
  int32_t* order = nullptr;    // cells sorted by level
  uint8_t* events = nullptr;
  int32_t* flag_scan = nullptr;  // scans (cap+1)
  int32_t* flag_scan2 = nullptr;
  int32_t* misc = nullptr;       // small counters
  int32_t* level_count = nullptr;
  uint32_t* claim = nullptr;     // per-voxel birth claims
  marl_cell_state* birth_cells = nullptr;
  marl_ruleset* birth_rules = nullptr;
  int32_t* birth_flag = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  int32_t* pinned = nullptr;   // small pinned host buffer
  double* budget = nullptr;    // device FP64 partial sums
  int kin_mode = MARL_KINETICS_PLAIN;
  SurrogateState* surr = nullptr;  // learned kinetics surrogate, created on first use

  // device-driven tick captured as a hipGraph
  void* tdev = nullptr;            // per-tick counters on the device (TickDev)
  void* tick_ring = nullptr;       // pinned host copies of a run of captured ticks' counters
  // one captured tick per field parity (a tick with an odd number of stencil
  // launches alternates between them)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<uint8_t> graph_key[2];  // inputs each captured graph was recorded for
  int graph_flip[2] = {0, 0};         // the graph swaps the field buffers
  int graph_sweeps = 64, graph_levels = 64;  // schedule budget of the captured tick
  bool graph_auto = true;                    // budget follows the schedule's needs
  int cell_layout = MARL_CELL_LAYOUT_AUTO;    // field layout of the plain cell pass
  bool kin_quads = false;                    // the layout last chosen: quads in place
  // which crossover marl_choose_cell_layout uses: a single-column tick, whose
  // last substep writes the view, or an ensemble column, which transposes both ways
  bool layout_view = true;
  bool graph_overlap = false;                // the captured tick overlaps phases on ctx->aux
  bool graph_small = false;                  // the captured cell phase runs in one workgroup
  int kin_block = 128;                       // threads per block of the plain pass (64, 128 or 256)
  long fault_calls = 0;                      // drop-in / tick calls counted for MARL_HIP_FAULT=tick:<n>
  bool phase_budget = false;                 // FP64 mass budget at the tick's phase boundaries
  double* phase_budget_dev = nullptr;        // [4][40] of the last tick
  int relax_sweeps = 4;          // sweeps the last eager schedule needed to converge
  marl_sim_params last_sim{};

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // A second stream for work that overlaps the tick's critical path (the
  // conflict schedule beside the field phase, the transpose back beside
  // fate), and untimed events to fork and join it.
  hipStream_t aux = nullptr;
  hipEvent_t xev[4] = {nullptr, nullptr, nullptr, nullptr};
};

// Runs the enclosing scope's launches on another stream: helpers that enqueue
// on ctx->stream then use that stream; the previous one is restored on exit.
struct OnStream {
  marl_ctx* c;
  hipStream_t prev;
  OnStream(marl_ctx* ctx, hipStream_t s) : c(ctx), prev(ctx->stream) { c->stream = s; }
  ~OnStream() { c->stream = prev; }
  OnStream(const OnStream&) = delete;
  OnStream& operator=(const OnStream&) = delete;
};

// helpers implemented in ctx.hip
int marl_fail(marl_ctx* ctx, int code, const std::string& msg);
int marl_hip_fail(marl_ctx* ctx, hipError_t e, const char* what);
#define MARL_HIP(ctx, expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return marl_hip_fail((ctx), _e, #expr); \
  } while (0)

// Allocation-free population summary (output.hip): the file-scope scratch
// array's size and the deepest grid it serves.
constexpr int kNoAllocSummaryMaxZ = 4096;
constexpr size_t kNoAllocScratchBytes = 72 * 1024;
// Looks up the scratch array's address on `device` (which must be current);
// called once per device by marl_create.
int marl_noalloc_init_device(int device);
// MARL_HIP_FAULT hooks (ctx.hip).
bool marl_fault_init();
int marl_fault_tick(marl_ctx* ctx);

StencilParams marl_stencil_params(const marl_sim_params* sim, int apply_sources);
int marl_check_params(marl_ctx* ctx, const marl_sim_params* sim);

// kernels launched from other translation units
int marl_launch_substep(marl_ctx* ctx, const StencilParams& p, int tile);
// n substeps of sim with the given tile (two per launch for temporal tiles),
// boundary sources fused into the first when sources_first; *launches (may
// be null) = kernel launches, i.e. field buffer swaps.
int marl_launch_substeps(marl_ctx* ctx, const marl_sim_params* sim, int n, bool sources_first, int tile,
                         int* launches = nullptr);
// A tick's field phase: boundary sources, the substeps and the light scan,
// with the sources fused into the first substep and, where the tile allows,
// the light scan into the last.
// With `view` (ctx->staging) the last substep writes the new field into that
// AoS view instead of the quads and ctx->view_fresh is set, so the cell pass
// skips its transpose in; a phase that cannot do so (no single last substep)
// writes the quads and clears the flag.
int marl_launch_field_phase(marl_ctx* ctx, const marl_sim_params* sim, bool sources, float* view = nullptr);
int marl_launch_boundary(marl_ctx* ctx, const marl_sim_params* sim);
int marl_launch_light(marl_ctx* ctx, const marl_sim_params* sim);
int marl_launch_light_compact(marl_ctx* ctx, const marl_sim_params* sim);
// Ensembles: fill the tile fields of `cols` for a tile id (returns total
// logical tiles and light blocks), and launch one batched substep / light scan.
// A column's z-chunk is its whole depth with whole_columns, else chunks[c]
// planes when chunks is given and chunks[c] > 0, else the tile shape's
// chunk, cut at the column's depth.
void marl_ens_layout(int tile, EnsColumn* cols, int ncols, int* total_tiles, int* total_light_blocks,
                     bool whole_columns, const int* chunks = nullptr);
// pieces: nullptr (the tile numbering of the layout) or a device copy of a
// marl_ens_pieces table for a persistent grid of `grid` blocks.
hipError_t marl_launch_substep_ens(hipStream_t st, const EnsColumn* dcols, int ncols, int total_tiles, int parity,
                                   const StencilParams& p, int tile, int grid, const uint32_t* pieces = nullptr);
// Piece table of the balanced persistent walk: words 0..8 are the first
// piece of each XCD's run (and the end), then 3 words per piece: column |
// species group << 16, tile x | tile y << 16, z-chunk index | chunk planes << 16.
constexpr int kEnsPieceHeader = 12;
void marl_ens_pieces(int tile, const EnsColumn* cols, int ncols, int grid, std::vector<uint32_t>& out);
hipError_t marl_launch_light_ens(hipStream_t st, const EnsColumn* dcols, int ncols, int total_blocks,
                                 const marl_sim_params* sim, bool compact);
int marl_stencil_ens_grid(int tile, int total_tiles);

// Device-driven tick pieces (cells.hip), shared by the single-column graph and
// ensembles.  prepare allocates; the enqueue calls are capturable.
int marl_graph_prepare(marl_ctx* ctx, const marl_sim_params* sim);
int marl_enqueue_tick_begin(marl_ctx* ctx, const marl_sim_params* sim);
int marl_enqueue_field_phase(marl_ctx* ctx, const marl_sim_params* sim);
int marl_enqueue_cell_phase(marl_ctx* ctx, const marl_sim_params* sim);
// The cell phase in two parts: the conflict schedule, which reads only the
// cells and the occupancy and so may run beside the field phase, and the
// ordered pass with fate and compaction after both (with fork_back the
// transpose back of the pass's field view forks onto ctx->aux and joins before
// the part ends).
int marl_enqueue_schedule(marl_ctx* ctx, const marl_sim_params* sim);
int marl_enqueue_cell_phase_small(marl_ctx* ctx, const marl_sim_params* sim);
int marl_enqueue_pass_and_fate(marl_ctx* ctx, const marl_sim_params* sim, bool fork_back);
std::vector<uint8_t> marl_graph_key(const marl_ctx* ctx, const marl_sim_params* sim);
int marl_graph_stage(marl_ctx* ctx, hipStream_t st, int64_t tick);
int marl_graph_fetch(marl_ctx* ctx, hipStream_t st);
// The same for many contexts with one launch each way (ensembles): the live
// counts and tick go to every column's device counters in one kernel; the
// counters come back gathered into `dev_out` (device, n x marl_tickdev_bytes())
// and copied to `host_out` (pinned) in one transfer; after the stream has
// synchronised, marl_graph_fetch_finish_many hands each column its copy.
size_t marl_tickdev_bytes();
int marl_graph_stage_many(marl_ctx* const* cols, int n, hipStream_t st, int64_t tick);
int marl_graph_fetch_many(marl_ctx* const* cols, int n, hipStream_t st, void* const* dev_srcs, void* dev_out,
                          void* host_out);
void marl_graph_fetch_finish_many(marl_ctx* const* cols, int n, const void* host_out);
int marl_graph_post(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick, marl_tick_stats* st,
                    int32_t* host_finished);
int marl_launch_occupancy_from_cells(marl_ctx* ctx);
int marl_ensure_cell_capacity(marl_ctx* ctx, int64_t need);
// AoS (ctx->staging) <-> resident quads (dst/src)
int marl_transpose(marl_ctx* ctx, bool to_quads, float* quads);
// species planes (ctx->staging) <-> resident quads
int marl_planes(marl_ctx* ctx, bool to_quads, float* quads);
int marl_pack_occ(marl_ctx* ctx);
int marl_unpack_occ(marl_ctx* ctx);
int marl_ensure_staging(marl_ctx* ctx);
int marl_ensure_aux(marl_ctx* ctx);
// Picks (and records in ctx->kin_quads) the plain cell pass's field layout.
bool marl_choose_cell_layout(marl_ctx* ctx);
int marl_launch_copy_probe(marl_ctx* ctx);
int marl_launch_mfma_probe(marl_ctx* ctx, int cus, float* out, double* flop);
size_t marl_mfma_probe_floats(int cus);
int marl_launch_gather_probe(marl_ctx* ctx, const void* src, uint32_t nsectors, float* out, size_t threads,
                             uint32_t salt, double* sector_bytes);

// stencil tiles: rows per thread and z-chunk
struct TileShape { int ry, zc, pd, sg, tb, nw; };  // tb: substeps per pass (2 = temporal blocking); nw: waves/block
extern const TileShape kTiles[];
extern const int kNumTiles;
bool marl_tile_ensemble_ok(int tile);
int marl_default_tile(const marl_dims& d);
