// ensemble.hip — ensembles of independent columns of uneven sizes and seeds,
// advanced together.
//
// No reference counterpart: the reference runs one column per process
// (crates/marl-sim/src/lib.rs:160-277).  Every column is an ordinary context
// (its own dims, cells and seed); the ensemble owns a stream and one hipGraph
// per tick shape:
//   - fork: each column's counters and occupancy on the column's stream;
//   - join, then the field phase of all columns in one persistent stencil
//     launch per substep and one light launch (k_stencil_ens, k_light_ens);
//   - fork: each column's cell phase (schedule, ordered pass, fate, compaction)
//     on its own stream, so small columns run side by side; join.
// Results are bit-identical to ticking each column alone.
#include <algorithm>
#include <chrono>
#include <functional>
#include <cstring>
#include <vector>

#include "marl_internal.h"

struct marl_ensemble {
  std::vector<marl_ctx*> cols;
  hipStream_t stream = nullptr;
  hipEvent_t fork[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> join;  // two per column
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<EnsColumn> hcols;
  EnsColumn* dcols = nullptr;
  int tile = 7;
  int per_tile_grid = 0;  // 0: persistent grid; 1: one block per stencil tile; 2: persistent, balanced pieces
  uint32_t* dpieces = nullptr;  // device piece table of grid mode 2
  size_t dpieces_cap = 0;       // its capacity in words
  int whole = 0;          // 1: every column's z-chunk is its whole depth and the light scan rides the last substep
  std::vector<int> chunks;  // per column: z-chunk planes (0: the tile shape's), when not whole
  int total_tiles = 0, total_light = 0, grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> key;
  std::string err;
  // batched counter transfers: the columns' device counter pointers, the
  // gathered counters on the device and their pinned host copy
  std::vector<void*> tdevs;
  void** d_tdevs = nullptr;
  void* d_counters = nullptr;
  void* h_counters = nullptr;
};

namespace {

int efail(marl_ensemble* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

#define ENS_HIP(e, call)                                                                            \
  do {                                                                                              \
    hipError_t _err = (call);                                                                       \
    if (_err != hipSuccess) return efail((e), MARL_ERR_TRANSFER, std::string(#call ": ") + hipGetErrorString(_err)); \
  } while (0)

// a column's failure, reported on the ensemble
int col_fail(marl_ensemble* e, int c, int rc) {
  return efail(e, rc, "column " + std::to_string(c) + ": " + marl_last_error(e->cols[c]));
}

void drop(marl_ensemble* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  e->key.clear();
}

// The batched kernels' view of the columns for the ensemble's tile shape and
// grid, on the host.
// The table lists the columns longest stencil tiles first (z-chunk, then
// row length), so that the first blocks of a launch, whether one block per
// tile or a persistent grid taking tiles round-robin, get the longest tiles
// and the shortest ones fill the end.  Only the tile numbering depends on the
// order; every column's arithmetic is the same.
// order (if given): table entry k holds column order[k].
void host_table(const marl_ensemble* e, std::vector<EnsColumn>& cols, int* total_tiles, int* total_light,
                int* grid, std::vector<int>* order = nullptr) {
  const int n = (int)e->cols.size();
  const bool has_chunks = (int)e->chunks.size() == n;
  auto chunk_of = [&](int c) {
    const int Z = e->cols[c]->d.z, zc = kTiles[e->tile].zc;
    if (e->whole) return Z;
    if (has_chunks && e->chunks[c] > 0) return std::min(e->chunks[c], Z);
    return std::min(zc, Z);
  };
  std::vector<int> ord(n);
  for (int c = 0; c < n; ++c) ord[c] = c;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    const int za = chunk_of(a), zb = chunk_of(b);
    if (za != zb) return za > zb;
    return e->cols[a]->d.x > e->cols[b]->d.x;
  });
  if (order) *order = ord;
  std::vector<int> chunks(n, 0);
  cols.assign(n, EnsColumn{});
  for (int k = 0; k < n; ++k) {
    const int c = ord[k];
    if (has_chunks) chunks[k] = e->chunks[c];
    marl_ctx* x = e->cols[c];
    EnsColumn& t = cols[k];
    t.f[0] = x->field[0];
    t.f[1] = x->field[1];
    t.occ = x->occ;
    t.light = x->light;
    t.X = x->d.x;
    t.Y = x->d.y;
    t.Z = x->d.z;
    t.XW = x->xw;
    t.cur = x->cur;
  }
  const bool pieces = e->per_tile_grid == 2;
  marl_ens_layout(e->tile, cols.data(), n, total_tiles, total_light, e->whole != 0 && !pieces,
                  has_chunks && !pieces ? chunks.data() : nullptr);
  if (pieces) {  // a multiple of 8 blocks, at least one per XCD
    const int g = marl_stencil_ens_grid(e->tile, 1 << 30);
    *grid = std::max(8, g / 8 * 8);
  } else {
    *grid = e->per_tile_grid ? *total_tiles : marl_stencil_ens_grid(e->tile, *total_tiles);
  }
}

// ... and its device copy.
int build_table(marl_ensemble* e) {
  const int n = (int)e->cols.size();
  host_table(e, e->hcols, &e->total_tiles, &e->total_light, &e->grid);
  if (!e->dcols) ENS_HIP(e, hipMalloc(&e->dcols, sizeof(EnsColumn) * n));
  ENS_HIP(e, hipMemcpyAsync(e->dcols, e->hcols.data(), sizeof(EnsColumn) * n, hipMemcpyHostToDevice, e->stream));
  if (e->per_tile_grid == 2) {
    std::vector<uint32_t> pieces;
    marl_ens_pieces(e->tile, e->hcols.data(), n, e->grid, pieces);
    if (pieces.size() > e->dpieces_cap) {
      ENS_HIP(e, hipStreamSynchronize(e->stream));
      if (e->dpieces) (void)hipFree(e->dpieces);
      e->dpieces = nullptr;
      e->dpieces_cap = 0;
      ENS_HIP(e, hipMalloc(&e->dpieces, pieces.size() * sizeof(uint32_t)));
      e->dpieces_cap = pieces.size();
    }
    ENS_HIP(e, hipMemcpyAsync(e->dpieces, pieces.data(), pieces.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));  // `pieces` must outlive the copy
  }
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  return MARL_OK;
}

// The piece table the batched stencil walks, or nullptr.
const uint32_t* launch_pieces(const marl_ensemble* e) { return e->per_tile_grid == 2 ? e->dpieces : nullptr; }

// The buffers of the batched counter transfers, rebuilt when a column's
// device counters moved (they are allocated by the column's first prepare).
int ensure_counter_buffers(marl_ensemble* e) {
  const int n = (int)e->cols.size();
  bool same = (int)e->tdevs.size() == n;
  for (int c = 0; same && c < n; ++c) same = e->tdevs[c] == e->cols[c]->tdev;
  if (same && e->d_tdevs) return MARL_OK;
  e->tdevs.resize(n);
  for (int c = 0; c < n; ++c) e->tdevs[c] = e->cols[c]->tdev;
  const size_t bytes = marl_tickdev_bytes() * (size_t)n;
  if (!e->d_tdevs) {
    ENS_HIP(e, hipMalloc((void**)&e->d_tdevs, sizeof(void*) * n));
    ENS_HIP(e, hipMalloc(&e->d_counters, bytes));
    ENS_HIP(e, hipHostMalloc(&e->h_counters, bytes, hipHostMallocDefault));
  }
  ENS_HIP(e, hipMemcpyAsync(e->d_tdevs, e->tdevs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, e->stream));
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  return MARL_OK;
}

std::vector<uint8_t> ens_key(marl_ensemble* e, const std::vector<marl_sim_params>& sims) {
  std::vector<uint8_t> k;
  for (size_t c = 0; c < e->cols.size(); ++c) {
    const std::vector<uint8_t> kc = marl_graph_key(e->cols[c], &sims[c]);
    k.insert(k.end(), kc.begin(), kc.end());
  }
  k.push_back((uint8_t)e->tile);
  k.push_back((uint8_t)e->per_tile_grid);
  k.push_back((uint8_t)e->whole);
  for (int z : e->chunks) {  // chunk lengths fit 16 bits (a field is under 2 GiB)
    k.push_back((uint8_t)(z & 0xff));
    k.push_back((uint8_t)((z >> 8) & 0xff));
  }
  return k;
}

// The light scan rides the last substep when every column is one z-chunk
// deep (whole columns; the balanced pieces may cut a column into chunks).
bool whole_light(const marl_ensemble* e) { return e->whole && e->per_tile_grid != 2; }

// The batched field phase of one tick: the substeps (sources fused into the
// first) and the light scan, fused into the last substep when every column is
// one z-chunk deep, else its own launch.  parity_fixed: every launch reads the
// current buffers (timing without changing the fields).
int enqueue_field_phase(marl_ensemble* e, const marl_sim_params& sim, bool parity_fixed) {
  const int n = (int)e->cols.size();
  const int S = sim.diffusion_substeps;
  for (int k = 0; k < S; ++k) {
    StencilParams p = marl_stencil_params(&sim, k == 0 ? 1 : 0);
    if (k == S - 1) {  // the light scan itself, or the compact species 4 the light launch reads
      p.fuse_light = whole_light(e) ? 1 : 2;
      p.l_i0 = sim.surface_intensity;
      p.l_e_cell = marl_expf(-sim.cell_absorption);
      p.l_nchem = -sim.chemical_absorption;
      p.l_floor = sim.light_floor;
    }
    ENS_HIP(e, marl_launch_substep_ens(e->stream, e->dcols, n, e->total_tiles, parity_fixed ? 0 : (k & 1), p,
                                       e->tile, e->grid, launch_pieces(e)));
  }
  if (!whole_light(e)) ENS_HIP(e, marl_launch_light_ens(e->stream, e->dcols, n, e->total_light, &sim, true));
  return MARL_OK;
}

int enqueue_ensemble_tick(marl_ensemble* e, const std::vector<marl_sim_params>& sims) {
  const int n = (int)e->cols.size();
  const marl_sim_params& sim = sims[0];
  ENS_HIP(e, hipEventRecord(e->fork[0], e->stream));
  for (int c = 0; c < n; ++c) {
    marl_ctx* x = e->cols[c];
    ENS_HIP(e, hipStreamWaitEvent(x->stream, e->fork[0], 0));
    if (int rc = marl_enqueue_tick_begin(x, &sims[c])) return col_fail(e, c, rc);
    ENS_HIP(e, hipEventRecord(e->join[2 * c], x->stream));
    ENS_HIP(e, hipStreamWaitEvent(e->stream, e->join[2 * c], 0));
  }
  if (int rc = enqueue_field_phase(e, sim, false)) return rc;
  ENS_HIP(e, hipEventRecord(e->fork[1], e->stream));
  for (int c = 0; c < n; ++c) {
    marl_ctx* x = e->cols[c];
    ENS_HIP(e, hipStreamWaitEvent(x->stream, e->fork[1], 0));
    if (int rc = marl_enqueue_cell_phase(x, &sims[c])) return col_fail(e, c, rc);
    ENS_HIP(e, hipEventRecord(e->join[2 * c + 1], x->stream));
    ENS_HIP(e, hipStreamWaitEvent(e->stream, e->join[2 * c + 1], 0));
  }
  return MARL_OK;
}

int capture(marl_ensemble* e, const std::vector<marl_sim_params>& sims) {
  drop(e);
  if (int rc = build_table(e)) return rc;
  ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_ensemble_tick(e, sims);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc != MARL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  ENS_HIP(e, ee);
  e->graph = g;
  ENS_HIP(e, hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
  e->key = ens_key(e, sims);
  return MARL_OK;
}

std::vector<marl_sim_params> column_params(const marl_ensemble* e, const marl_sim_params* sim, const uint64_t* seeds) {
  std::vector<marl_sim_params> sims(e->cols.size(), *sim);
  for (size_t c = 0; c < sims.size(); ++c) sims[c].rng_seed = seeds ? seeds[c] : sim->rng_seed + c;
  return sims;
}

}  // namespace

extern "C" {

int marl_ensemble_create(marl_ctx* const* columns, int32_t n_columns, marl_ensemble** out) {
  if (!out) return MARL_ERR_INVALID_INPUT;
  *out = nullptr;
  if (!columns || n_columns < 1 || n_columns > 1024) return MARL_ERR_INVALID_INPUT;
  for (int c = 0; c < n_columns; ++c) {
    if (!columns[c]) return MARL_ERR_INVALID_INPUT;
    for (int d = 0; d < c; ++d)
      if (columns[d] == columns[c]) return MARL_ERR_INVALID_INPUT;
  }
  marl_ensemble* e = new marl_ensemble();
  e->cols.assign(columns, columns + n_columns);
  auto fail = [&](hipError_t) {
    marl_ensemble_destroy(e);
    return MARL_ERR_DEVICE_INIT;
  };
  hipError_t r;
  if ((r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) return fail(r);
  for (auto& f : e->fork)
    if ((r = hipEventCreateWithFlags(&f, hipEventDisableTiming)) != hipSuccess) return fail(r);
  e->join.assign(2 * n_columns, nullptr);
  for (auto& j : e->join)
    if ((r = hipEventCreateWithFlags(&j, hipEventDisableTiming)) != hipSuccess) return fail(r);
  if ((r = hipEventCreate(&e->ev0)) != hipSuccess) return fail(r);
  if ((r = hipEventCreate(&e->ev1)) != hipSuccess) return fail(r);
  *out = e;
  return MARL_OK;
}

void marl_ensemble_destroy(marl_ensemble* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop(e);
  if (e->dcols) (void)hipFree(e->dcols);
  if (e->dpieces) (void)hipFree(e->dpieces);
  if (e->d_tdevs) (void)hipFree(e->d_tdevs);
  if (e->d_counters) (void)hipFree(e->d_counters);
  if (e->h_counters) (void)hipHostFree(e->h_counters);
  for (auto f : e->fork)
    if (f) (void)hipEventDestroy(f);
  for (auto j : e->join)
    if (j) (void)hipEventDestroy(j);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* marl_ensemble_last_error(const marl_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int marl_ensemble_set_tile(marl_ensemble* e, int32_t tile) {
  if (!e) return MARL_ERR_INVALID_INPUT;
  if (tile < 0 || tile >= marl_stencil_tile_count()) return efail(e, MARL_ERR_INVALID_INPUT, "unknown tile id");
  if (!marl_tile_ensemble_ok(tile))
    return efail(e, MARL_ERR_INVALID_INPUT, "ensembles use single-substep tiles");
  if (tile != e->tile) drop(e);
  e->tile = tile;
  return MARL_OK;
}

int marl_ensemble_set_grid(marl_ensemble* e, int32_t per_tile) {
  if (!e) return MARL_ERR_INVALID_INPUT;
  if (per_tile < 0 || per_tile > 2) return efail(e, MARL_ERR_INVALID_INPUT, "grid mode is 0, 1 or 2");
  if (per_tile != e->per_tile_grid) drop(e);
  e->per_tile_grid = per_tile;
  return MARL_OK;
}

int marl_ensemble_tick(marl_ensemble* e, const marl_sim_params* sim, const uint64_t* seeds, int64_t tick0,
                       int32_t n_ticks, marl_tick_stats* stats, int32_t* host_finished, float* ms_out) {
  if (!e || !sim) return MARL_ERR_INVALID_INPUT;
  if (n_ticks < 0) return efail(e, MARL_ERR_INVALID_INPUT, "negative tick count");
  if (sim->diffusion_substeps < 2 || sim->diffusion_substeps % 2 != 0)
    return efail(e, MARL_ERR_INVALID_INPUT, "ensembles need an even, positive substep count");
  const int n = (int)e->cols.size();
  const std::vector<marl_sim_params> sims = column_params(e, sim, seeds);
  int32_t finished = 0;
  float ms_total = 0.0f;
  std::vector<marl_tick_stats> st(n);
  for (int t = 0; t < n_ticks; ++t) {
    for (int c = 0; c < n; ++c) {
      e->cols[c]->layout_view = false;  // the batched field phase writes quads only
      if (int rc = marl_graph_prepare(e->cols[c], &sims[c])) return col_fail(e, c, rc);
    }
    if (e->key != ens_key(e, sims))
      if (int rc = capture(e, sims)) return rc;
    if (int rc = ensure_counter_buffers(e)) return rc;
    // every column's live count and tick in one launch, every column's
    // counters back in one gather and one copy
    if (int rc = marl_graph_stage_many(e->cols.data(), n, e->stream, tick0 + t)) return col_fail(e, 0, rc);
    ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
    ENS_HIP(e, hipGraphLaunch(e->gexec, e->stream));
    ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
    if (int rc = marl_graph_fetch_many(e->cols.data(), n, e->stream, e->d_tdevs, e->d_counters, e->h_counters))
      return col_fail(e, 0, rc);
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    marl_graph_fetch_finish_many(e->cols.data(), n, e->h_counters);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
    ms_total += ms;
    for (int c = 0; c < n; ++c) {
      if (int rc = marl_graph_post(e->cols[c], &sims[c], tick0 + t, &st[c], &finished)) return col_fail(e, c, rc);
      st[c].ms_field = ms;
    }
  }
  if (stats)
    for (int c = 0; c < n; ++c) stats[c] = st[c];
  if (host_finished) *host_finished = finished;
  if (ms_out) *ms_out = ms_total;
  return MARL_OK;
}

int marl_ensemble_time_field(marl_ensemble* e, const marl_sim_params* sim, int32_t n_substeps, float* ms_out) {
  if (!e || !sim || n_substeps <= 0 || !ms_out) return MARL_ERR_INVALID_INPUT;
  drop(e);  // the table is rebuilt for the current tile
  if (int rc = build_table(e)) return rc;
  const int n = (int)e->cols.size();
  const StencilParams sp = marl_stencil_params(sim, 0);
  // every launch reads each column's current buffer and writes the other one,
  // so the resident fields are left unchanged
  ENS_HIP(e, marl_launch_substep_ens(e->stream, e->dcols, n, e->total_tiles, 0, sp, e->tile, e->grid,
                                     launch_pieces(e)));
  ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
  for (int k = 0; k < n_substeps; ++k)
    ENS_HIP(e, marl_launch_substep_ens(e->stream, e->dcols, n, e->total_tiles, 0, sp, e->tile, e->grid,
                                       launch_pieces(e)));
  ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
  ENS_HIP(e, hipEventSynchronize(e->ev1));
  float ms = 0.0f;
  ENS_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  *ms_out = ms / (float)n_substeps;
  return MARL_OK;
}

namespace {

// Two timed field phases of the current launch shape after a warm-up, summed
// (every launch reads the current buffers; the light planes are rewritten).
int time_phase(marl_ensemble* e, const marl_sim_params& sim, float* ms_out) {
  drop(e);
  if (int rc = build_table(e)) return rc;
  float total = 0.0f;
  for (int rep = 0; rep < 3; ++rep) {  // the first is a warm-up
    ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
    if (int rc = enqueue_field_phase(e, sim, true)) return rc;
    ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
    ENS_HIP(e, hipEventSynchronize(e->ev1));
    float ms = 0.0f;
    ENS_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    if (rep > 0) total += ms;
  }
  *ms_out = total;
  return MARL_OK;
}

}  // namespace

// The ensemble's tile shape and grid (persistent or one block per tile),
// picked by timing every candidate on the resident columns (the fields are
// left unchanged), then the z-chunk of each group of columns of one depth.
int marl_ensemble_autotune(marl_ensemble* e, const marl_sim_params* sim, int32_t* tile_out) {
  if (!e || !sim) return MARL_ERR_INVALID_INPUT;
  if (sim->diffusion_substeps <= 0) return efail(e, MARL_ERR_INVALID_INPUT, "autotune needs substeps");
  const int n = (int)e->cols.size();
  // the fused light writes the columns' light planes: keep copies to restore
  std::vector<float*> saved(n, nullptr);
  struct Free {
    std::vector<float*>& p;
    ~Free() {
      for (float* q : p)
        if (q) (void)hipFree(q);
    }
  } fr{saved};
  for (int c = 0; c < n; ++c) {
    const size_t bytes = e->cols[c]->nvox * sizeof(float);
    ENS_HIP(e, hipMalloc(&saved[c], bytes));
    ENS_HIP(e, hipMemcpyAsync(saved[c], e->cols[c]->light, bytes, hipMemcpyDeviceToDevice, e->stream));
  }
  // Every tile shape, grid mode and chunk mode with the tile shape's chunks.
  struct Cand {
    float ms;
    int tile, grid, whole;
  };
  std::vector<Cand> cands;
  e->chunks.clear();
  for (int t = 0; t < marl_stencil_tile_count(); ++t) {
    if (!marl_tile_ensemble_ok(t)) continue;
    for (int g = 0; g < 3; ++g)
      for (int w = 0; w < (g == 2 ? 1 : 2); ++w) {
        e->tile = t;
        e->per_tile_grid = g;
        e->whole = w;
        float ms = 0.0f;
        if (int rc = time_phase(e, *sim, &ms)) return rc;
        cands.push_back({ms, t, g, w});
      }
  }
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
  Cand best = cands.front();
  std::vector<int> best_chunks(n, 0);
  // Then, for the five fastest shapes without whole columns, the z-chunk of
  // each group of equally deep columns (deepest first, two passes): a
  // column's chunk trades the halo planes every chunk re-reads against how
  // evenly the tiles fill the chip.
  std::vector<int> depths;
  for (marl_ctx* x : e->cols) depths.push_back(x->d.z);
  std::sort(depths.begin(), depths.end(), std::greater<int>());
  depths.erase(std::unique(depths.begin(), depths.end()), depths.end());
  const int opts[] = {4, 8, 16, 24, 32, 48, 64};
  int tried = 0;
  for (const Cand& c0 : cands) {
    if (c0.whole || c0.grid == 2 || tried == 5) continue;  // the pieces set their own chunks
    ++tried;
    e->tile = c0.tile;
    e->per_tile_grid = c0.grid;
    e->whole = 0;
    e->chunks.assign(n, 0);
    float cur_ms = c0.ms;
    for (int pass = 0; pass < 2; ++pass)
    for (int depth : depths) {
      int keep_zc = 0;
      for (int c = 0; c < n; ++c)  // the second pass starts from the first pass's choice
        if (e->cols[c]->d.z == depth) keep_zc = e->chunks[c];
      for (int k = -1; k < (int)(sizeof(opts) / sizeof(opts[0])); ++k) {
        const int zc = k < 0 ? depth : opts[k];  // k = -1: one chunk per column
        if (k >= 0 && zc >= depth) continue;
        for (int c = 0; c < n; ++c)
          if (e->cols[c]->d.z == depth) e->chunks[c] = zc;
        float ms = 0.0f;
        if (int rc = time_phase(e, *sim, &ms)) return rc;
        if (ms < cur_ms) {
          cur_ms = ms;
          keep_zc = zc;
        }
      }
      for (int c = 0; c < n; ++c)
        if (e->cols[c]->d.z == depth) e->chunks[c] = keep_zc;
    }
    if (cur_ms < best.ms) {
      best = {cur_ms, c0.tile, c0.grid, 0};
      best_chunks = e->chunks;
    }
  }
  e->tile = best.tile;
  e->per_tile_grid = best.grid;
  e->whole = best.whole;
  e->chunks = best.whole ? std::vector<int>(n, 0) : best_chunks;
  for (int c = 0; c < n; ++c)
    ENS_HIP(e, hipMemcpyAsync(e->cols[c]->light, saved[c], e->cols[c]->nvox * sizeof(float),
                              hipMemcpyDeviceToDevice, e->stream));
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  // each column's plain-pass block size on its own population (the column's
  // state is restored after every trial pass)
  for (int c = 0; c < n; ++c)
    if (int rc = marl_autotune_kinetics(e->cols[c], sim, nullptr)) return col_fail(e, c, rc);
  drop(e);
  if (tile_out) *tile_out = best.tile;
  return MARL_OK;
}

int marl_ensemble_time_field_phase(marl_ensemble* e, const marl_sim_params* sim, int32_t n_ticks, float* ms_out) {
  if (!e || !sim || n_ticks <= 0 || !ms_out || sim->diffusion_substeps <= 0) return MARL_ERR_INVALID_INPUT;
  drop(e);
  if (int rc = build_table(e)) return rc;
  if (int rc = enqueue_field_phase(e, *sim, true)) return rc;  // warm-up
  ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
  for (int t = 0; t < n_ticks; ++t)
    if (int rc = enqueue_field_phase(e, *sim, true)) return rc;
  ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
  ENS_HIP(e, hipEventSynchronize(e->ev1));
  float ms = 0.0f;
  ENS_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  *ms_out = ms / (float)n_ticks;
  return MARL_OK;
}

int marl_ensemble_set_whole_columns(marl_ensemble* e, int32_t whole) {
  if (!e) return MARL_ERR_INVALID_INPUT;
  e->whole = whole != 0;
  drop(e);
  return MARL_OK;
}

int marl_ensemble_set_chunks(marl_ensemble* e, const int32_t* chunks, int32_t n) {
  if (!e) return MARL_ERR_INVALID_INPUT;
  const int nc = (int)e->cols.size();
  if (n != 0 && (n != nc || !chunks))
    return efail(e, MARL_ERR_INVALID_INPUT, "marl_ensemble_set_chunks: one entry per column, or n = 0");
  for (int c = 0; c < n; ++c)
    if (chunks[c] < 0 || chunks[c] > 65535) return efail(e, MARL_ERR_INVALID_INPUT, "z-chunk out of range");
  if (n == 0) e->chunks.assign(nc, 0);
  else e->chunks.assign(chunks, chunks + n);
  drop(e);
  return MARL_OK;
}

int marl_ensemble_get_chunks(const marl_ensemble* e, int32_t* chunks, int32_t n) {
  if (!e || !chunks || n != (int)e->cols.size()) return MARL_ERR_INVALID_INPUT;
  std::vector<EnsColumn> cols;
  std::vector<int> order;
  int total = 0, light = 0, grid = 0;
  host_table(e, cols, &total, &light, &grid, &order);
  for (int k = 0; k < n; ++k) chunks[order[k]] = cols[k].zc;
  return MARL_OK;
}

int marl_ensemble_launch_shape(const marl_ensemble* e, int32_t* tile, int32_t* grid_blocks, int32_t* stencil_tiles,
                               int32_t* whole_columns) {
  if (!e) return MARL_ERR_INVALID_INPUT;
  if (whole_columns) *whole_columns = whole_light(e);
  std::vector<EnsColumn> cols;
  int total = 0, light = 0, grid = 0;
  host_table(e, cols, &total, &light, &grid);
  if (e->per_tile_grid == 2) {  // the units of the walk are the pieces
    std::vector<uint32_t> pieces;
    marl_ens_pieces(e->tile, cols.data(), (int)cols.size(), grid, pieces);
    total = (int)pieces[8];
  }
  if (tile) *tile = e->tile;
  if (grid_blocks) *grid_blocks = grid;
  if (stencil_tiles) *stencil_tiles = total;
  return MARL_OK;
}

int marl_ensemble_get_grid(const marl_ensemble* e, int32_t* mode) {
  if (!e || !mode) return MARL_ERR_INVALID_INPUT;
  *mode = e->per_tile_grid;
  return MARL_OK;
}

}  // extern "C"
