// cells.hip — the GPU-resident cell path and the full tick.
//
// Reference: crates/marl-sim/src/lib.rs:160-277 (tick loop), spatial.rs:8-154
// (neighbour helpers, division placement), crates/marl-cell/src/cell.rs:193-524
// (CellState::tick, Ruleset::mutate).
//
// Exact sequential semantics of the cell pass (lib.rs:206-217: cells in Vec
// order, each one's deltas land in the field before the next reads it) on a
// data-parallel machine: two cells interact only through an empty voxel that
// is a face neighbour of both, i.e. when they are L1 distance 2 apart and a
// shared face neighbour is free.  Occupancy is fixed during the pass, so the
// conflict graph is fixed; level(j) = 1 + max level(i) over conflicting i < j.
// Running one launch per level, cells within a level touch disjoint voxels,
// and every voxel sees its writers in index order — the same result as the
// sequential loop, bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kinetics.h"
#include "marl_internal.h"
#include "surrogate.h"

namespace {

constexpr int kMaxLevels = 4096;
constexpr int kMaxPreds = 18;

CellParams cell_params(const marl_ctx* ctx, const marl_sim_params* s, uint64_t tick) {
  CellParams p;
  p.dt = s->dt;
  p.epsilon = s->epsilon;
  p.c_max = s->c_max;
  p.lambda_maintenance = s->lambda_maintenance;
  p.hard_death_floor = s->hard_death_floor;
  p.reaction_maintenance = s->reaction_maintenance;
  p.base_division_prep = s->base_division_prep;
  p.prep_maintenance_multiplier = s->prep_maintenance_multiplier;
  p.rush_penalty_rate = s->rush_penalty_rate;
  p.active_reaction_threshold = s->active_reaction_threshold;
  p.mutation_stddev = s->mutation_stddev;
  p.structural_mutation_rate_mult = s->structural_mutation_rate_mult;
  p.meta_mutation_rate = s->meta_mutation_rate;
  p.meta_lo = s->meta_mutation_clamp_low;
  p.meta_hi = s->meta_mutation_clamp_high;
  p.hill_lo = s->hill_exponent_clamp_low;
  p.hill_hi = s->hill_exponent_clamp_high;
  p.div_dist = s->division_neighbor_distance;
  p.X = ctx->d.x;
  p.Y = ctx->d.y;
  p.Z = ctx->d.z;
  p.seed = s->rng_seed;
  p.stream = (MARL_TAG_DIVISION << 32) | (tick & 0xffffffffull);
  p.fplane = (uint64_t)ctx->nvox * 4;
  return p;
}

// FACE_OFFSETS (spatial.rs:8-15): +x, -x, +y, -y, +z, -z
__constant__ int8_t kFace[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// Conflict candidates at L1 distance 2 (the cells that can share a free face
// neighbour with a cell): +-2 along each axis in face order, then one step
// along each pair of faces (a, b), a < b, on different axes.  kCandFaces: the
// faces of this cell that lie between it and the candidate (the midpoint, or
// the two corners).
constexpr int kCand = 18;
__constant__ int8_t kCandOff[kCand][3] = {
    {2, 0, 0},  {-2, 0, 0}, {0, 2, 0},  {0, -2, 0}, {0, 0, 2},  {0, 0, -2},                          // axes
    {1, 1, 0},  {1, -1, 0}, {1, 0, 1},  {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1},  // x with y, z
    {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};                                                  // y with z
__constant__ uint8_t kCandFaces[kCand] = {
    1, 2, 4, 8, 16, 32,
    1 | 4, 1 | 8, 1 | 16, 1 | 32, 2 | 4, 2 | 8, 2 | 16, 2 | 32,
    4 | 16, 4 | 32, 8 | 16, 8 | 32};

__device__ inline bool inb(int x, int y, int z, const CellParams& p) {
  return x >= 0 && x < p.X && y >= 0 && y < p.Y && z >= 0 && z < p.Z;
}
__device__ inline int64_t vox(int x, int y, int z, const CellParams& p) {
  return ((int64_t)z * p.Y + y) * p.X + x;
}

// ---------------------------------------------------------------------------
// utility kernels
__global__ void k_fill_i32(int32_t* a, int32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) a[i] = v;
}

// K3: occupancy bits and the voxel -> cell index map (≙ cell_map + the
// occupancy Vec<bool> of lib.rs:174-179).
__global__ void k_scatter_cells(const marl_cell_state* __restrict__ cells, int n, int32_t* __restrict__ cell_at,
                                uint32_t* __restrict__ occ, CellParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint16_t* ps = cells[i].pos;
  const int64_t v = vox(ps[0], ps[1], ps[2], p);
  cell_at[v] = i;
  atomicOr(&occ[v >> 5], 1u << (v & 31));
}

// K4: empty-neighbour mask (spatial.rs:18-43) and the conflict predecessors.
// Conflicting partner at L1 distance 2 with a lower index that shares at least
// one free face neighbour.  Predecessor q of cell j is preds[q * pstride + j]
// (pstride = capacity): the relaxation sweeps then read predecessor lists row
// by row, coalesced across the cells of a wave.
__device__ __forceinline__ void cell_prep_cell(int j, const marl_cell_state* __restrict__ cells,
                                               const int32_t* __restrict__ cell_at, const uint32_t* __restrict__ occ,
                                               uint8_t* __restrict__ nbmask, int32_t* __restrict__ preds, size_t pstride,
                                               uint8_t* __restrict__ npred, int32_t* __restrict__ level,
                                               const CellParams& p) {
  const uint16_t* ps = cells[j].pos;
  const int x = ps[0], y = ps[1], z = ps[2];
  // Emptiness from the 1-bit mask (L2-resident), cell_at only for occupied
  // candidates.  Every mask word (6 faces, 18 candidates) is requested before
  // any is used, and the candidates' cell_at loads likewise, so a cell waits
  // for two rounds of loads, not one per candidate; out-of-grid positions read
  // word 0 and are masked off.
  uint32_t fw[6], fb[6];
  bool fin[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int nx = x + kFace[k][0], ny = y + kFace[k][1], nz = z + kFace[k][2];
    fin[k] = inb(nx, ny, nz, p);
    const int64_t w = fin[k] ? vox(nx, ny, nz, p) : 0;
    fw[k] = occ[w >> 5];
    fb[k] = (uint32_t)(w & 31);
  }
  int64_t cv[kCand];
  uint32_t cw[kCand];
  bool cin[kCand];
#pragma unroll
  for (int c = 0; c < kCand; ++c) {
    const int ox = x + kCandOff[c][0], oy = y + kCandOff[c][1], oz = z + kCandOff[c][2];
    cin[c] = inb(ox, oy, oz, p);
    cv[c] = cin[c] ? vox(ox, oy, oz, p) : 0;
    cw[c] = occ[cv[c] >> 5];
  }
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (fin[k] && !((fw[k] >> fb[k]) & 1u)) m |= 1u << k;
  nbmask[j] = (uint8_t)m;
  // a candidate conflicts when it is occupied and one of the faces it shares
  // with this cell (kCandFaces) is free
  int32_t ci[kCand];
#pragma unroll
  for (int c = 0; c < kCand; ++c) {
    const bool want = cin[c] && ((cw[c] >> (cv[c] & 31)) & 1u) && (m & kCandFaces[c]);
    ci[c] = want ? cell_at[cv[c]] : -1;
  }
  int np = 0;
  int32_t* pj = preds + j;
#pragma unroll
  for (int c = 0; c < kCand; ++c)
    if (ci[c] >= 0 && ci[c] < j) pj[(size_t)np++ * pstride] = ci[c];
  npred[j] = (uint8_t)np;
  level[j] = 0;
}

__global__ void k_cell_prep(const marl_cell_state* __restrict__ cells, int n, const int32_t* __restrict__ cell_at,
                            const uint32_t* __restrict__ occ, uint8_t* __restrict__ nbmask,
                            int32_t* __restrict__ preds, size_t ps, uint8_t* __restrict__ npred,
                            int32_t* __restrict__ level, CellParams p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cell_prep_cell(j, cells, cell_at, occ, nbmask, preds, ps, npred, level, p);
}

// Longest-path relaxation on the conflict DAG; monotone, converges to the
// exact level in (depth + 1) sweeps or fewer.
// Sweep k (0-based) of the monotone relaxation.  Levels only rise and never
// pass the longest-path depth; after k sweeps every level is at least
// min(depth, k) (the synchronous bound, which in-place updates only beat), so
// a level below k is already the depth and the cell is skipped without
// reading its predecessors' levels: the sweeps' scattered reads shrink with
// the frontier instead of costing every cell every sweep.
__device__ __forceinline__ bool relax_settled(const int32_t* level, int j, int k) { return level[j] < k; }

__global__ void k_level_relax(int k, int n, const int32_t* __restrict__ preds, size_t ps,
                              const uint8_t* __restrict__ npred, int32_t* level, int32_t* changed) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (relax_settled(level, j, k)) return;
  const int np = npred[j];
  if (np == 0) return;
  const int32_t* pj = preds + j;
  int L = 0;
  for (int k = 0; k < np; ++k) L = max(L, level[pj[(size_t)k * ps]] + 1);
  if (L > level[j]) {
    level[j] = L;
    *changed = 1;
  }
}

// Level histogram and counting-sort scatter.  Most cells share a few levels,
// so per-lane atomics on those counters would serialize: lanes of a wave with
// the same level are grouped by ballot (one atomic per group, in LDS), blocks
// aggregate in LDS, and one global atomic per block and level remains.
constexpr int kLocalLevels = 64;

// Groups the lanes of a wave by `key` (all 64 lanes must call it).  For each
// distinct key among the valid lanes, the group's first lane calls
// per_group(key, count); each valid lane returns that call's result plus its
// rank among the group's lanes (-1 for invalid lanes).
template <typename F>
__device__ __forceinline__ int wave_group(bool valid, int key, F&& per_group) {
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long active = __ballot(valid);
  int out = -1;
  while (active) {
    const int ld = __ffsll(active) - 1;
    const int kl = __shfl(key, ld);
    const unsigned long long m = __ballot(valid && key == kl);
    int b = 0;
    if (lane == ld) b = per_group(kl, __popcll(m));
    b = __shfl(b, ld);
    if (valid && key == kl) out = b + __popcll(m & ((1ull << lane) - 1ull));
    active &= ~m;
  }
  return out;
}

constexpr int kLevelBlocks = 256;  // blocks of the histogram and scatter: few same-address atomics

__device__ __forceinline__ void level_hist_block(int n, const int32_t* __restrict__ level,
                                                 int32_t* hist /* [0]=max, [1+L]=count */) {
  __shared__ int32_t lh[kLocalLevels];
  __shared__ int32_t lmax;
  for (int t = threadIdx.x; t < kLocalLevels; t += blockDim.x) lh[t] = 0;
  if (threadIdx.x == 0) lmax = 0;
  __syncthreads();
  for (int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {  // uniform per block
    const int j = b0 + threadIdx.x;
    const bool valid = j < n;
    wave_group(valid, valid ? level[j] : 0, [&](int L, int c) {
      atomicMax(&lmax, L);
      if (L < kLocalLevels) atomicAdd(&lh[L], c);
      else if (L < kMaxLevels) atomicAdd(&hist[1 + L], c);
      return 0;
    });
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kLocalLevels; t += blockDim.x)
    if (lh[t]) atomicAdd(&hist[1 + t], lh[t]);
  if (threadIdx.x == 0) atomicMax(&hist[0], lmax);
}

// Counting-sort scatter: block b owns a contiguous run of cells, counts its
// levels in LDS, reserves one range per level with a single global atomic,
// then places its cells.
__device__ __forceinline__ void level_scatter_block(int n, const int32_t* __restrict__ level, int32_t* cursor,
                                                    int32_t* __restrict__ order) {
  __shared__ int32_t cnt[kLocalLevels], base[kLocalLevels];
  const int chunk = ((n + 255) / 256 + (int)gridDim.x - 1) / (int)gridDim.x * 256;
  const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  for (int t = threadIdx.x; t < kLocalLevels; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  // the levels of 8 sub-chunks are loaded together, so their latencies overlap
  constexpr int U = 8;
  for (int c0 = lo; c0 < hi; c0 += U * 256) {
    int Ls[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = c0 + u * 256 + (int)threadIdx.x;
      Ls[u] = j < hi ? level[j] : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      wave_group(Ls[u] >= 0 && Ls[u] < kLocalLevels, Ls[u], [&](int k, int c) { return atomicAdd(&cnt[k], c); });
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kLocalLevels; t += blockDim.x) {
    base[t] = cnt[t] ? atomicAdd(&cursor[t], cnt[t]) : 0;
    cnt[t] = 0;  // reused as the running count of the placement pass
  }
  __syncthreads();
  for (int c0 = lo; c0 < hi; c0 += U * 256) {
    int Ls[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = c0 + u * 256 + (int)threadIdx.x;
      Ls[u] = j < hi ? level[j] : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = c0 + u * 256 + (int)threadIdx.x;
      // deep levels (beyond the LDS counters) take their global slots directly
      const int pos = wave_group(Ls[u] >= 0, Ls[u], [&](int k, int c) {
        return k < kLocalLevels ? base[k] + atomicAdd(&cnt[k], c) : atomicAdd(&cursor[k], c);
      });
      if (Ls[u] >= 0) order[pos] = j;
    }
  }
}

__global__ void __launch_bounds__(256) k_level_hist(int n, const int32_t* __restrict__ level, int32_t* hist) {
  level_hist_block(n, level, hist);
}

__global__ void __launch_bounds__(256) k_level_scatter(int n, const int32_t* __restrict__ level, int32_t* cursor,
                                                       int32_t* __restrict__ order) {
  level_scatter_block(n, level, cursor, order);
}

// ---------------------------------------------------------------------------
// K5/K6: one level of the cell pass.  Per cell: mean environment of the free
// face neighbours (spatial.rs:56-77), CellState::tick (cell.rs:193-369), then
// split deltas with the max(0) clamp into those neighbours (spatial.rs:87-110,
// field.rs:57-62).  The field is read and written through an AoS view (one
// 48-byte record per voxel, transposed around the pass) so a neighbour costs
// three float4 accesses in one or two cache lines; with QUADS (sparse
// populations, where the two whole-field transposes cost more than the extra
// lines) the resident species quads are used in place.  Dynamically
// indexed per-cell arrays live in LDS (one private slot per thread, odd
// strides), so nothing goes to scratch.  With a harvest sink, sampled cells
// also append their (features, post-kinetics deltas and fluxes) pair.
constexpr int kKinBlock = 128;  // default block of the plain pass; marl_autotune_kinetics picks 64, 128 or 256

// The six face records gathered for the environment are kept for the
// scatter, which then writes without reading them again.
template <bool HARVEST, bool QUADS>
__device__ __forceinline__ void kinetics_cell(int i, marl_cell_state* __restrict__ cells,
                                              const marl_ruleset* __restrict__ rules, float* __restrict__ field,
                                              const float* __restrict__ light, const uint8_t* __restrict__ nbmask,
                                              uint8_t* __restrict__ events, const CellParams& p,
                                              const HarvestSink& hv, float* sI, float* sE, float* sD, float* sQ,
                                              float* sR, float* sG, float* sH) {
  marl_cell_state c = load_cell(cells + i);
  const marl_ruleset& r = rules[i];
  const uint32_t m = nbmask[i];
  const size_t v = ((size_t)c.pos[2] * p.Y + c.pos[1]) * p.X + c.pos[0];
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) sI[k] = c.internal[k];
  float4 keep[18];
  gather_env6_keep<QUADS>(field, v, m, p, sE, keep);
  const float lt = light[v];
  PtrArr I{sI}, D{sD};
  const PtrArr E{sE};
  const AosRules ar{&r};
  // a sampled cell of the harvest also records the explicit step from its pre-tick state
  const bool take = HARVEST && (uint64_t)sample_hash(hv.tick, (uint32_t)i) < hv.threshold;
  float frel[MARL_R_MAX];
  if (take) {
    PtrArr dI{sQ}, dD{sR}, gI{sG}, gD{sH};
    explicit_step(c, ar, PtrArr{sI}, E, dI, dD, lt, p, frel, gI, gD);
  }
  cell_kinetics(c, ar, I, E, D, lt, p);
  const float4 fate = ar.fate();
  // one counter update per wave: a per-cell atomic would serialise on one address
  unsigned long long slot = 0;
  if (HARVEST) {
    const unsigned long long mask = __ballot(take);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (mask && lane == leader) base = atomicAdd(hv.count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader);
    slot = base + __popcll(mask & ((1ull << lane) - 1ull));
  }
  if (take) {
    if ((int64_t)slot < hv.cap) {
      float f[kFeat];
      make_features(c.internal, c, sE, lt, sQ, sR, frel, f);
      float* fo = hv.feat + slot * kFeat;
#pragma unroll
      for (int k = 0; k < kFeat; ++k) fo[k] = f[k];
      float* to = hv.targ + slot * kTarg;
      float sc[kTarg];
      target_scales(sG, sH, sc);
#pragma unroll
      for (int k = 0; k < MARL_M_INT - 1; ++k) to[k] = scaled_target((sI[k] - c.internal[k]) - sQ[k], sc[k]);
      to[MARL_M_INT - 1] = 0.0f;
#pragma unroll
      for (int s2 = 0; s2 < MARL_S_EXT; ++s2) to[16 + s2] = scaled_target(sD[s2] - sR[s2], sc[16 + s2]);
      hv.hash[slot] = topology_hash(r);
    }
  }
  const uint8_t ev = cell_fate(c, fate, sI[0], p);
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) c.internal[k] = sI[k];
  cells[i] = c;
  events[i] = ev;
  scatter_from_kept<QUADS>(field, v, m, p, sD, keep);
}

// One level: `order` points at the level's first cell.
template <bool HARVEST, bool QUADS, int B = kKinBlock>
__global__ void __launch_bounds__(B) k_kinetics(const int32_t* __restrict__ order, int count,
                                              marl_cell_state* __restrict__ cells,
                                              const marl_ruleset* __restrict__ rules, float* __restrict__ field,
                                              const float* __restrict__ light, const uint8_t* __restrict__ nbmask,
                                              uint8_t* __restrict__ events, CellParams p, HarvestSink hv) {
  __shared__ float sI[B * 17];
  __shared__ float sE[B * 13];
  __shared__ float sD[B * 13];
  __shared__ float sQ[HARVEST ? B * 17 : 1];  // the harvest's explicit-step deltas
  __shared__ float sR[HARVEST ? B * 13 : 1];
  __shared__ float sG[HARVEST ? B * 17 : 1];  // and their gross activities
  __shared__ float sH[HARVEST ? B * 13 : 1];
  const int t = blockIdx.x * B + threadIdx.x;
  if (t >= count) return;
  const int i = order[t];
  kinetics_cell<HARVEST, QUADS>(i, cells, rules, field, light, nbmask, events, p, hv, sI + threadIdx.x * 17,
                                sE + threadIdx.x * 13, sD + threadIdx.x * 13, HARVEST ? sQ + threadIdx.x * 17 : sQ,
                                HARVEST ? sR + threadIdx.x * 13 : sR, HARVEST ? sG + threadIdx.x * 17 : sG,
                                HARVEST ? sH + threadIdx.x * 13 : sH);
}

// ---------------------------------------------------------------------------
// K7: fate.  Division targets (find_empty_neighbor, spatial.rs:115-154) in the
// pre-birth map; one RNG stream per dividing cell keyed (seed, tick, index).
__device__ __forceinline__ void div_target_cell(int i, const marl_cell_state* __restrict__ cells,
                                                const uint8_t* __restrict__ events,
                                                const int32_t* __restrict__ cell_at, int32_t* __restrict__ target,
                                                int32_t* __restrict__ bflag, const CellParams& p) {
  int32_t tgt = -1;
  if (events[i] == MARL_EVENT_DIVISION) {
    const uint16_t* ps = cells[i].pos;
    const int x = ps[0], y = ps[1], z = ps[2];
    int32_t cand[6];
    int nc = 0;
    const int dd = p.div_dist;
    for (int k = 0; k < 6; ++k) {
      const int nx = x + kFace[k][0] * dd, ny = y + kFace[k][1] * dd, nz = z + kFace[k][2] * dd;
      if (inb(nx, ny, nz, p)) {
        const int64_t w = vox(nx, ny, nz, p);
        if (cell_at[w] < 0) cand[nc++] = (int32_t)w;
      }
    }
    if (nc == 0) {
      for (int k = 0; k < 6; ++k) {
        const int nx = x + kFace[k][0], ny = y + kFace[k][1], nz = z + kFace[k][2];
        if (inb(nx, ny, nz, p)) {
          const int64_t w = vox(nx, ny, nz, p);
          if (cell_at[w] < 0) cand[nc++] = (int32_t)w;
        }
      }
    }
    if (nc > 0) {
      MarlRng rng;
      rng.init(p.seed, p.stream, (uint64_t)i);
      tgt = cand[rng.range((uint32_t)nc)];
    }
  }
  target[i] = tgt;
  bflag[i] = tgt >= 0 ? 1 : 0;
}

__global__ void k_div_target(int n, const marl_cell_state* __restrict__ cells, const uint8_t* __restrict__ events,
                             const int32_t* __restrict__ cell_at, int32_t* __restrict__ target,
                             int32_t* __restrict__ bflag, CellParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  div_target_cell(i, cells, events, cell_at, target, bflag, p);
}

// Ruleset::mutate (cell.rs:406-524) in its draw order, on the daughter's copy.
__device__ inline void maybe_mutate(float& v, float rate, float sd, MarlRng& rng) {
  if (rng.u01() < rate) {
    v += rng.normal(sd);
    v = fmaxf(v, 0.0f);
  }
}

__device__ void mutate_ruleset(marl_ruleset* r, const CellParams& p, MarlRng& rng) {
  const float rate = r->mutation_rate;
  const float sd = p.mutation_stddev;
  const float sr = rate * p.structural_mutation_rate_mult;
  for (int q = 0; q < MARL_S_RECEPTORS; ++q) {
    marl_receptor* rc = &r->receptors[q];
    maybe_mutate(rc->k_half, rate, sd, rng);
    maybe_mutate(rc->n_hill, rate, sd, rng);
    rc->n_hill = clamp_rs(rc->n_hill, p.hill_lo, p.hill_hi);
    maybe_mutate(rc->gain, rate, sd, rng);
  }
  for (int q = 0; q < MARL_S_TRANSPORTERS; ++q) {
    marl_transport* t = &r->transport[q];
    maybe_mutate(t->uptake_rate, rate, sd, rng);
    maybe_mutate(t->secrete_rate, rate, sd, rng);
    if (rng.u01() < sr) t->ext_species = (uint8_t)rng.range(MARL_S_EXT);
    if (rng.u01() < sr) t->int_species = (uint8_t)rng.range(MARL_M_INT);
  }
  int active[MARL_R_MAX];
  int na = 0;
  for (int q = 0; q < MARL_R_MAX; ++q)
    if (fabsf(r->reactions[q].v_max) > p.active_reaction_threshold) active[na++] = q;
  for (int q = 0; q < MARL_R_MAX; ++q) {
    marl_reaction* rx = &r->reactions[q];
    maybe_mutate(rx->k_m, rate, sd, rng);
    maybe_mutate(rx->v_max, rate, sd, rng);
    maybe_mutate(rx->k_cat, rate, sd, rng);
    for (int fld = 0; fld < 3; ++fld) {
      if (rng.u01() < sr) {
        const uint32_t idx = rng.range((uint32_t)(na > 1 ? na : 1));
        if ((int)idx < na) {
          const int donor = active[idx];
          if (donor != q) {
            if (fld == 0) rx->substrate = r->reactions[donor].substrate;
            else if (fld == 1) rx->product = r->reactions[donor].product;
            else rx->catalyst = r->reactions[donor].catalyst;
          }
        } else {
          const uint8_t vv = (uint8_t)rng.range(MARL_M_INT);
          if (fld == 0) rx->substrate = vv;
          else if (fld == 1) rx->product = vv;
          else rx->catalyst = vv;
        }
      }
    }
  }
  for (int q = 0; q < MARL_S_EFFECTORS; ++q) {
    maybe_mutate(r->effectors[q].threshold, rate, sd, rng);
    maybe_mutate(r->effectors[q].rate, rate, sd, rng);
  }
  maybe_mutate(r->division_energy, rate, sd, rng);
  maybe_mutate(r->death_energy, rate, sd, rng);
  maybe_mutate(r->quiescence_energy, rate, sd, rng);
  maybe_mutate(r->division_prep_ticks, rate, sd, rng);
  r->division_prep_ticks = fmaxf(r->division_prep_ticks, 1.0f);
  maybe_mutate(r->hgt_propensity, rate, sd, rng);
  if (rng.u01() < p.meta_mutation_rate) {
    r->mutation_rate += rng.normal(sd) * p.meta_mutation_rate;
    r->mutation_rate = clamp_rs(r->mutation_rate, p.meta_lo, p.meta_hi);
  }
}

// Births (lib.rs:226-243): daughter = clone of the ticked parent at the target,
// age 0, prep 0, new lineage; both halve all 16 internals; daughter mutates.
// Birth slot = rank of the division in event order; claims keep the first
// birth per voxel (lib.rs:270-277).
__device__ __forceinline__ void birth_cell(int i, marl_cell_state* __restrict__ cells,
                                           const marl_ruleset* __restrict__ rules,
                                           const int32_t* __restrict__ target, const int32_t* __restrict__ brank,
                                           marl_cell_state* __restrict__ bcells, marl_ruleset* __restrict__ brules,
                                           int32_t* __restrict__ btarget, uint32_t* __restrict__ claim,
                                           const CellParams& p, const uint64_t* tab = nullptr, int tab_n = 0) {
  const int32_t tgt = target[i];
  if (tgt < 0) return;
  const int32_t rk = brank[i];
  MarlRng rng;
  rng.init(p.seed, p.stream, (uint64_t)i, 1);  // draw 0 picked the target
  rng.tab = tab;
  rng.tab_n = (uint32_t)tab_n;
  marl_cell_state d = cells[i];
  const int64_t XY = (int64_t)p.X * p.Y;
  d.pos[0] = (uint16_t)(tgt % p.X);
  d.pos[1] = (uint16_t)((tgt / p.X) % p.Y);
  d.pos[2] = (uint16_t)(tgt / XY);
  d.age = 0;
  d.prep_remaining = 0;
  d.lineage_id = rng.u64();
  marl_cell_state par = cells[i];
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) {
    d.internal[k] *= 0.5f;
    par.internal[k] *= 0.5f;
  }
  cells[i] = par;
  bcells[rk] = d;
  marl_ruleset* dr = &brules[rk];
  *dr = rules[i];
  mutate_ruleset(dr, p, rng);
  btarget[rk] = tgt;
  atomicMin(&claim[tgt], (uint32_t)rk);
}

// A division's daughter draws some two hundred numbers in a data-dependent
// order (mutate's draws depend on earlier outcomes), one Philox block per
// eight.  One-wave blocks look at 64 cells at a time, find the dividing ones
// by ballot, and for each its 64 lanes compute the cell's first 64 blocks
// together (draws 0-511, more than a mutation uses), then lane 0 walks the
// draws from that table; the stream and its order are unchanged.
constexpr int kBirthBlocks = 64;  // one per lane
constexpr int kBirthGrid = 4096;  // one-wave blocks of the births kernels

__device__ __forceinline__ void birth_cell_wave(int i, marl_cell_state* __restrict__ cells,
                                                const marl_ruleset* __restrict__ rules,
                                                const int32_t* __restrict__ target, const int32_t* __restrict__ brank,
                                                marl_cell_state* __restrict__ bcells, marl_ruleset* __restrict__ brules,
                                                int32_t* __restrict__ btarget, uint32_t* __restrict__ claim,
                                                const CellParams& p, uint64_t* tab) {
  const int lane = (int)threadIdx.x;  // i is the same in every lane: one cell per block of one wave
  uint64_t c[4] = {(uint64_t)lane, (uint64_t)i, 0, 0};
  marl_philox4x64(c, p.seed, p.stream);
#pragma unroll
  for (int k = 0; k < 4; ++k) tab[4 * lane + k] = c[k];
  __syncthreads();
  if (lane == 0) birth_cell(i, cells, rules, target, brank, bcells, brules, btarget, claim, p, tab, kBirthBlocks);
  __syncthreads();  // the table is reused for the block's next cell
}

__device__ __forceinline__ void births_scan(int n, marl_cell_state* __restrict__ cells,
                                            const marl_ruleset* __restrict__ rules,
                                            const int32_t* __restrict__ target, const int32_t* __restrict__ brank,
                                            marl_cell_state* __restrict__ bcells, marl_ruleset* __restrict__ brules,
                                            int32_t* __restrict__ btarget, uint32_t* __restrict__ claim,
                                            const CellParams& p, uint64_t* tab) {
  // block b takes cells b, b + G, b + 2G, ... (G blocks), 64 of them per
  // ballot: dividing cells, which come in runs of neighbouring indices, are
  // spread over the blocks
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  for (int64_t k0 = 0; b + k0 * G < n; k0 += 64) {
    const int64_t i = b + (k0 + (int64_t)threadIdx.x) * G;
    unsigned long long m = __ballot(i < n && target[i] >= 0);
    while (m) {
      const int l = __ffsll((long long)m) - 1;
      m &= m - 1;
      birth_cell_wave((int)(b + (k0 + l) * G), cells, rules, target, brank, bcells, brules, btarget, claim, p, tab);
    }
  }
}

__global__ void __launch_bounds__(64) k_births(int n, marl_cell_state* __restrict__ cells,
                                               const marl_ruleset* __restrict__ rules,
                                               const int32_t* __restrict__ target, const int32_t* __restrict__ brank,
                                               marl_cell_state* __restrict__ bcells, marl_ruleset* __restrict__ brules,
                                               int32_t* __restrict__ btarget, uint32_t* __restrict__ claim,
                                               CellParams p) {
  __shared__ uint64_t tab[kBirthBlocks * 4];
  births_scan(n, cells, rules, target, brank, bcells, brules, btarget, claim, p, tab);
}

__global__ void k_birth_winners(int nb, const int32_t* __restrict__ btarget, const uint32_t* __restrict__ claim,
                                int32_t* __restrict__ win) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  win[r] = claim[btarget[r]] == (uint32_t)r ? 1 : 0;
}

__global__ void k_claim_reset(int nb, const int32_t* __restrict__ btarget, uint32_t* __restrict__ claim) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  claim[btarget[r]] = 0xffffffffu;
}

__global__ void k_death_flags(int n, const uint8_t* __restrict__ events, int32_t* __restrict__ dflag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dflag[i] = events[i] == MARL_EVENT_DEATH ? 1 : 0;
}

// K8: swap_remove compaction (lib.rs:257-267) in closed form.  With P the
// exclusive prefix count of deaths, level j covers [lo_j, n) with head end
// m_j = n - (P[n] - P[lo_j]) and lo_{j+1} = m_j.  Output slot s < m_0 holds
// s itself if it survives; otherwise the q-th smallest head hole of level j
// (q = P[s] - P[lo_j]) receives what lands in slot m_j + q of level j+1.
// Survivors below m_0 keep their slot, so only the head holes are written:
// each one's source (a tail survivor at or past m_0) is found here and the
// records are moved in place by k_move_records.
__device__ __forceinline__ void hole_source(int s0, int n, int m0, const uint8_t* __restrict__ events,
                                            const int32_t* __restrict__ P, int32_t* __restrict__ hole_dst,
                                            int32_t* __restrict__ hole_src) {
  if (events[s0] != MARL_EVENT_DEATH) return;
  const int total = P[n];
  int s = s0, lo = 0, m = m0;
  while (events[s] == MARL_EVENT_DEATH) {
    const int q = P[s] - P[lo];
    const int nlo = m;
    s = m + q;
    lo = nlo;
    m = n - (total - P[lo]);
  }
  const int h = P[s0];  // every death before s0 is a head hole too
  hole_dst[h] = s0;
  hole_src[h] = s;
}

__global__ void k_hole_sources(int n, int m0, const uint8_t* __restrict__ events, const int32_t* __restrict__ P,
                               int32_t* __restrict__ hole_dst, int32_t* __restrict__ hole_src) {
  const int s0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (s0 >= m0) return;
  hole_source(s0, n, m0, events, P, hole_dst, hole_src);
}

// Moves cell + ruleset records in 16-byte chunks, one chunk per lane, so a
// record's 42 chunks are copied by consecutive lanes.
constexpr int kCellChunks = sizeof(marl_cell_state) / 16;
constexpr int kRuleChunks = sizeof(marl_ruleset) / 16;
constexpr int kRecChunks = kCellChunks + kRuleChunks;
static_assert(sizeof(marl_cell_state) % 16 == 0 && sizeof(marl_ruleset) % 16 == 0, "record sizes");

__device__ __forceinline__ void move_chunk(int64_t t, const int32_t* __restrict__ dst_idx,
                                           const int32_t* __restrict__ src_idx,
                                           const marl_cell_state* __restrict__ src_c,
                                           const marl_ruleset* __restrict__ src_r, marl_cell_state* __restrict__ dst_c,
                                           marl_ruleset* __restrict__ dst_r, const int32_t* __restrict__ select) {
  const int h = (int)(t / kRecChunks), c = (int)(t % kRecChunks);
  if (select && !select[h]) return;
  const int d = dst_idx ? dst_idx[h] : h;
  const int sidx = src_idx ? src_idx[h] : h;
  if (c < kCellChunks) {
    reinterpret_cast<uint4*>(dst_c + d)[c] = reinterpret_cast<const uint4*>(src_c + sidx)[c];
  } else {
    reinterpret_cast<uint4*>(dst_r + d)[c - kCellChunks] = reinterpret_cast<const uint4*>(src_r + sidx)[c - kCellChunks];
  }
}

__global__ void k_move_records(int count, const int32_t* __restrict__ dst_idx, const int32_t* __restrict__ src_idx,
                               const marl_cell_state* __restrict__ src_c, const marl_ruleset* __restrict__ src_r,
                               marl_cell_state* __restrict__ dst_c, marl_ruleset* __restrict__ dst_r,
                               const int32_t* __restrict__ select) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)count * kRecChunks) return;
  move_chunk(t, dst_idx, src_idx, src_c, src_r, dst_c, dst_r, select);
}

// Destination slot of each surviving birth: m0 + its rank among the winners.
__global__ void k_birth_slots(int nb, int base, const int32_t* __restrict__ win, const int32_t* __restrict__ wrank,
                              int32_t* __restrict__ slot) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  slot[r] = win[r] ? base + wrank[r] : 0;
}

// K9: FP64 mass budget, deterministic two-pass reduction.
constexpr int kBudget = 40;
__global__ void __launch_bounds__(256) k_budget_field(const float* __restrict__ f, const uint32_t* __restrict__ occ,
                                                      size_t nvox, double* __restrict__ partial) {
  double acc[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) acc[k] = 0.0;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    const int o = (occ[v >> 5] >> (v & 31)) & 1u;
    const float4* f4 = reinterpret_cast<const float4*>(f) + v;
    const float4 q[3] = {f4[0], f4[nvox], f4[2 * nvox]};
#pragma unroll
    for (int s = 0; s < MARL_S_EXT; ++s) {
      const float4& a = q[s >> 2];
      const double x = (s & 3) == 0 ? a.x : (s & 3) == 1 ? a.y : (s & 3) == 2 ? a.z : a.w;
      acc[s] += o ? 0.0 : x;
      acc[12 + s] += o ? x : 0.0;
    }
  }
  __shared__ double red[256];
  for (int k = 0; k < 24; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * kBudget + k] = red[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_budget_cells(const marl_cell_state* __restrict__ c, int n,
                                                      double* __restrict__ partial) {
  double acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] += (double)c[i].internal[k];
  __shared__ double red[256];
  for (int k = 0; k < 16; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * kBudget + 24 + k] = red[0];
    __syncthreads();
  }
}

__global__ void k_budget_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= kBudget) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * kBudget + k];
  out[k] = s;
}

// ---------------------------------------------------------------------------
// Device-driven tick.  The same per-cell functions as above, but every count
// (live cells, schedule depth, births, deaths, holes) is read from a small
// device record instead of the host, kernels loop over their work with a grid
// stride, and scans run over the buffer capacity.  A whole tick then needs no
// host round trip and is captured once as a hipGraph and replayed.  The
// schedule gets a fixed budget of relaxation sweeps and level launches (spare
// ones exit at once); a tick that would exceed it, or the cell capacity, sets
// `err`, skips the cell phase and is finished by the host-driven path.
constexpr int kGraphSweeps = 64;
constexpr int64_t kGraphOverlapCells = 131072;  // captured ticks overlap phases from this population up
constexpr int kGraphLevels = 64;  // <= kLocalLevels: every level is block-aggregated
static_assert(kGraphLevels <= kLocalLevels, "graph levels");

struct TickDev {
  int32_t n, err, nlev, nb, nd, nw, m0, holes;
  int32_t divisions, deaths, births_added, levels;
  uint64_t tick;
  int32_t chg[kGraphSweeps];
  int32_t off[kGraphLevels + 1];
};

__device__ __forceinline__ uint64_t div_stream(const TickDev* ts) {
  return (MARL_TAG_DIVISION << 32) | (ts->tick & 0xffffffffull);
}

#define GS_LOOP(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
                           i += (int64_t)gridDim.x * blockDim.x)

__global__ void k_g_begin(TickDev* ts, int64_t cap) {
  if (threadIdx.x == 0) {
    ts->err = (2 * (int64_t)ts->n + 16 > cap) ? 1 : 0;
    ts->nlev = ts->nb = ts->nd = ts->nw = ts->m0 = ts->holes = 0;
    ts->divisions = ts->deaths = ts->births_added = ts->levels = 0;
  }
  for (int k = threadIdx.x; k < kGraphSweeps; k += blockDim.x) ts->chg[k] = 0;
}

__global__ void k_g_scatter_cells(const marl_cell_state* __restrict__ cells, const TickDev* ts,
                                  int32_t* __restrict__ cell_at, uint32_t* __restrict__ occ, CellParams p) {
  const int n = ts->n;
  GS_LOOP(i, n) {
    const uint16_t* ps = cells[i].pos;
    const int64_t v = vox(ps[0], ps[1], ps[2], p);
    cell_at[v] = (int32_t)i;
    atomicOr(&occ[v >> 5], 1u << (v & 31));
  }
}

__global__ void k_g_cell_prep(const marl_cell_state* __restrict__ cells, const TickDev* ts,
                              const int32_t* __restrict__ cell_at, const uint32_t* __restrict__ occ,
                              uint8_t* __restrict__ nbmask, int32_t* __restrict__ preds, size_t ps,
                              uint8_t* __restrict__ npred, int32_t* __restrict__ level, CellParams p) {
  if (ts->err) return;
  const int n = ts->n;
  GS_LOOP(j, n) cell_prep_cell((int)j, cells, cell_at, occ, nbmask, preds, ps, npred, level, p);
}

// sweep k runs only if sweep k-1 changed a level
__global__ void k_g_level_relax(int k, TickDev* ts, const int32_t* __restrict__ preds, size_t ps,
                                const uint8_t* __restrict__ npred, int32_t* level) {
  if (ts->err || (k > 0 && ts->chg[k - 1] == 0)) return;
  const int n = ts->n;
  bool any = false;
  GS_LOOP(j, n) {
    if (relax_settled(level, (int)j, k)) continue;
    const int np = npred[j];
    if (np == 0) continue;
    const int32_t* pj = preds + j;
    int L = 0;
    for (int q = 0; q < np; ++q) L = max(L, level[pj[(size_t)q * ps]] + 1);
    if (L > level[j]) {
      level[j] = L;
      any = true;
    }
  }
  if (any) ts->chg[k] = 1;
}

__global__ void k_g_relax_check(TickDev* ts, int sweeps) {
  if (ts->err == 0 && ts->chg[sweeps - 1] != 0) ts->err = 2;  // not known to have converged
}

__global__ void __launch_bounds__(256) k_g_level_hist(const TickDev* ts, const int32_t* __restrict__ level,
                                                      int32_t* hist) {
  if (ts->err) return;
  level_hist_block(ts->n, level, hist);
}

__global__ void k_g_level_offsets(TickDev* ts, const int32_t* __restrict__ hist, int32_t* __restrict__ cursor,
                                  int max_levels) {
  if (threadIdx.x != 0 || ts->err) return;
  const int n = ts->n;
  const int nlev = n > 0 ? hist[0] + 1 : 0;
  if (nlev > max_levels) {
    ts->err = 3;
    return;
  }
  int acc = 0;
  for (int L = 0; L < nlev; ++L) {
    ts->off[L] = acc;
    cursor[L] = acc;
    acc += hist[1 + L];
  }
  ts->off[nlev] = acc;
  ts->nlev = nlev;
  ts->levels = nlev;
}

// Levels here are below kLocalLevels: deeper schedules set err before this kernel.
__global__ void __launch_bounds__(256) k_g_level_scatter(const TickDev* ts, const int32_t* __restrict__ level,
                                                         int32_t* cursor, int32_t* __restrict__ order) {
  if (ts->err) return;
  level_scatter_block(ts->n, level, cursor, order);
}

template <bool QUADS, int B>
__global__ void __launch_bounds__(B) k_g_kinetics(int L, const TickDev* ts, const int32_t* __restrict__ order,
                                                marl_cell_state* __restrict__ cells,
                                                const marl_ruleset* __restrict__ rules, float* __restrict__ field,
                                                const float* __restrict__ light, const uint8_t* __restrict__ nbmask,
                                                uint8_t* __restrict__ events, CellParams p) {
  if (ts->err || L >= ts->nlev) return;
  __shared__ float sI[B * 17];
  __shared__ float sE[B * 13];
  __shared__ float sD[B * 13];
  const int lo = ts->off[L], cnt = ts->off[L + 1] - lo;
  const HarvestSink hv{};
  GS_LOOP(t, cnt)
  kinetics_cell<false, QUADS>(order[lo + t], cells, rules, field, light, nbmask, events, p, hv,
                              sI + threadIdx.x * 17, sE + threadIdx.x * 13, sD + threadIdx.x * 13, nullptr, nullptr,
                              nullptr, nullptr);
}

// division targets for i < n; zero flags up to cap for the fixed-size scan
__global__ void k_g_div_target(const TickDev* ts, int64_t cap, const marl_cell_state* __restrict__ cells,
                               const uint8_t* __restrict__ events, const int32_t* __restrict__ cell_at,
                               int32_t* __restrict__ target, int32_t* __restrict__ bflag, CellParams p) {
  const int n = ts->err ? 0 : ts->n;
  p.stream = div_stream(ts);
  GS_LOOP(i, cap + 1) {
    if (i < n) {
      div_target_cell((int)i, cells, events, cell_at, target, bflag, p);
    } else {
      target[i] = -1;
      bflag[i] = 0;
    }
  }
}

__global__ void __launch_bounds__(64) k_g_births(const TickDev* ts, marl_cell_state* __restrict__ cells,
                                                 const marl_ruleset* __restrict__ rules,
                                                 const int32_t* __restrict__ target,
                                                 const int32_t* __restrict__ brank, marl_cell_state* __restrict__ bcells,
                                                 marl_ruleset* __restrict__ brules, int32_t* __restrict__ btarget,
                                                 uint32_t* __restrict__ claim, CellParams p) {
  __shared__ uint64_t tab[kBirthBlocks * 4];
  if (ts->err) return;
  const int n = ts->n;
  p.stream = div_stream(ts);
  births_scan(n, cells, rules, target, brank, bcells, brules, btarget, claim, p, tab);
}

__global__ void k_g_death_flags(const TickDev* ts, int64_t cap, const uint8_t* __restrict__ events,
                                int32_t* __restrict__ dflag) {
  const int n = ts->err ? 0 : ts->n;
  GS_LOOP(i, cap + 1) dflag[i] = (i < n && events[i] == MARL_EVENT_DEATH) ? 1 : 0;
}

__global__ void k_g_counts(TickDev* ts, const int32_t* __restrict__ brank, const int32_t* __restrict__ P) {
  if (threadIdx.x != 0 || ts->err) return;
  const int n = ts->n;
  const int nb = brank[n], nd = P[n], m0 = n - nd;
  ts->nb = nb;
  ts->nd = nd;
  ts->m0 = m0;
  ts->holes = (nd > 0 && m0 > 0) ? P[m0] : 0;
}

__global__ void k_g_birth_winners(const TickDev* ts, int64_t cap, const int32_t* __restrict__ btarget,
                                  const uint32_t* __restrict__ claim, int32_t* __restrict__ win) {
  const int nb = ts->err ? 0 : ts->nb;
  GS_LOOP(r, cap + 1) win[r] = (r < nb && claim[btarget[r]] == (uint32_t)r) ? 1 : 0;
}

__global__ void k_g_claim_reset(const TickDev* ts, const int32_t* __restrict__ btarget,
                                uint32_t* __restrict__ claim) {
  if (ts->err) return;
  const int nb = ts->nb;
  GS_LOOP(r, nb) claim[btarget[r]] = 0xffffffffu;
}

__global__ void k_g_hole_sources(TickDev* ts, const uint8_t* __restrict__ events, const int32_t* __restrict__ P,
                                 const int32_t* __restrict__ wrank, int32_t* __restrict__ hole_dst,
                                 int32_t* __restrict__ hole_src) {
  if (ts->err) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ts->nw = wrank[ts->nb];
  if (ts->holes == 0) return;
  const int n = ts->n, m0 = ts->m0;
  GS_LOOP(s0, m0) hole_source((int)s0, n, m0, events, P, hole_dst, hole_src);
}

__global__ void k_g_move_holes(const TickDev* ts, const int32_t* __restrict__ hole_dst,
                               const int32_t* __restrict__ hole_src, marl_cell_state* cells, marl_ruleset* rules) {
  if (ts->err) return;
  const int64_t total = (int64_t)ts->holes * kRecChunks;
  GS_LOOP(t, total) move_chunk(t, hole_dst, hole_src, cells, rules, cells, rules, nullptr);
}

__global__ void k_g_birth_slots(const TickDev* ts, const int32_t* __restrict__ win,
                                const int32_t* __restrict__ wrank, int32_t* __restrict__ slot) {
  if (ts->err) return;
  const int nb = ts->nb, m0 = ts->m0;
  GS_LOOP(r, nb) slot[r] = win[r] ? m0 + wrank[r] : 0;
}

__global__ void k_g_move_births(const TickDev* ts, const int32_t* __restrict__ slot,
                                const marl_cell_state* __restrict__ bcells, const marl_ruleset* __restrict__ brules,
                                marl_cell_state* __restrict__ cells, marl_ruleset* __restrict__ rules,
                                const int32_t* __restrict__ win) {
  if (ts->err || ts->nw == 0) return;
  const int64_t total = (int64_t)ts->nb * kRecChunks;
  GS_LOOP(t, total) move_chunk(t, slot, nullptr, bcells, brules, cells, rules, win);
}

__global__ void k_g_finish(TickDev* ts) {
  if (threadIdx.x != 0) return;
  if (ts->err == 0) {
    ts->divisions = ts->nb;
    ts->deaths = ts->nd;
    ts->births_added = ts->nw;
    ts->n = ts->m0 + ts->nw;
  }
  ts->tick += 1;
}

// ---------------------------------------------------------------------------
// The cell phase of a small population in one workgroup.  For a few hundred
// cells the device-driven phase above is some forty launches of a few
// microseconds each, most of them over almost no work; here one block of
// kSmallThreads runs the same steps with barriers between them: the masks and
// predecessors (cell_prep_cell), the relaxation to its fixed point with the
// levels in LDS, a counting sort of the cells by level, the levels in order
// (kinetics_cell on the resident quads), then fate, births with mutation,
// deaths and compaction with block scans over LDS.  Per cell it calls the
// same functions in the same order as the multi-launch phase, so the results
// are bit-identical.  Only a population above kSmallCells sets `err` (the host
// finishes the tick): the relaxation runs to its fixed point (at most n + 1
// sweeps) and every level is kept, however deep, so a run of small ticks can be
// queued without the host looking at each.  Global writes are read back only
// by this block after a barrier, so workgroup scope orders them.
constexpr int kSmallThreads = 512;
constexpr int kSmallCells = 512;
constexpr int kSmallRun = 16;  // captured small ticks queued per host look

// out[k] = in[0] + ... + in[k-1] for k in [0, m]; in[m] need not exist.
// All threads of the block call it; s_w holds kSmallThreads / 64 + 1 ints.
__device__ void block_exscan(const int32_t* in, int32_t* out, int m, int32_t* s_w) {
  constexpr int B = kSmallThreads, W = B / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int chunk = (m + 1 + B - 1) / B;
  const int lo = tid * chunk, hi = min(m, lo + chunk);  // out covers [lo, lo + chunk) within [0, m]
  int sum = 0;
  for (int k = lo; k < hi; ++k) sum += in[k];
  int v = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d);
    if (lane >= d) v += y;
  }
  if (lane == 63) s_w[wave] = v;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int w = 0; w < W; ++w) {
      const int t = s_w[w];
      s_w[w] = run;
      run += t;
    }
  }
  __syncthreads();
  int run = s_w[wave] + v - sum;
  const int top = min(m + 1, lo + chunk);
  for (int k = lo; k < top; ++k) {
    out[k] = run;
    if (k < m) run += in[k];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kSmallThreads) k_g_cell_phase_small(
    TickDev* ts, marl_cell_state* __restrict__ cells, marl_ruleset* __restrict__ rules, float* __restrict__ field,
    const float* __restrict__ light, const int32_t* __restrict__ cell_at, const uint32_t* __restrict__ occ,
    uint8_t* __restrict__ nbmask, int32_t* __restrict__ preds, size_t ps, uint8_t* __restrict__ npred,
    uint8_t* __restrict__ events, marl_cell_state* __restrict__ bcells, marl_ruleset* __restrict__ brules,
    uint32_t* __restrict__ claim, CellParams p) {
  constexpr int B = kSmallThreads, NC = kSmallCells;
  __shared__ float sI[B * 17];
  __shared__ float sE[B * 13];
  __shared__ float sD[B * 13];
  __shared__ int32_t s_level[NC + 1];        // levels; then the birth flags' ranks (brank)
  __shared__ int32_t s_order[NC + 1];        // cells by level; then the death flags
  __shared__ int32_t s_target[NC + 1];       // division targets; then the hole sources
  __shared__ int32_t s_flag[NC + 1];         // birth flags; then the winners
  __shared__ int32_t s_P[NC + 1];            // death ranks
  __shared__ int32_t s_btarget[NC + 1];      // each birth's voxel; then the hole destinations
  __shared__ int32_t s_wrank[NC + 1];        // winner ranks; then the birth slots
  __shared__ int32_t s_cnt[NC + 1], s_cur[NC + 1];  // per level (depth <= n)
  __shared__ int32_t s_w[B / 64 + 1];
  __shared__ int32_t s_flagv, s_depth;
  const int tid = threadIdx.x;
  if (ts->err) return;  // uniform: the whole block leaves
  const int n = ts->n;
  if (n > NC) {
    if (tid == 0) ts->err = 1;
    return;
  }
  // masks and predecessors (levels start at 0)
  for (int j = tid; j < n; j += B) cell_prep_cell(j, cells, cell_at, occ, nbmask, preds, ps, npred, s_level, p);
  __syncthreads();
  // relaxation to the fixed point (levels stay below n, so sweep n changes
  // nothing); sweep k skips levels below k (settled)
  int quiet = -1;
  for (int k = 0; k <= n; ++k) {
    if (tid == 0) s_flagv = 0;
    __syncthreads();
    for (int j = tid; j < n; j += B) {
      const int lj = s_level[j];
      const int np = npred[j];
      if (lj < k || np == 0) continue;
      int L = 0;
      for (int q = 0; q < np; ++q) L = max(L, s_level[preds[(size_t)q * ps + j]] + 1);
      if (L > lj) {
        s_level[j] = L;
        s_flagv = 1;
      }
    }
    __syncthreads();
    const int changed = s_flagv;
    if (tid == 0 && k < kGraphSweeps) ts->chg[k] = changed;
    __syncthreads();
    if (!changed) {
      quiet = k;
      break;
    }
  }
  if (quiet < 0) {  // not reached: a level never exceeds n - 1
    if (tid == 0) ts->err = 1;
    return;
  }
  // counting sort by level
  if (tid == 0) s_depth = 0;
  for (int L = tid; L <= NC; L += B) s_cnt[L] = 0;
  __syncthreads();
  for (int j = tid; j < n; j += B) {
    const int L = s_level[j];
    atomicMax(&s_depth, L + 1);
    atomicAdd(&s_cnt[L], 1);
  }
  __syncthreads();
  const int depth = s_depth;
  if (tid == 0) {
    int run = 0;
    for (int L = 0; L < depth; ++L) {
      const int c = s_cnt[L];
      s_cnt[L] = run;
      s_cur[L] = run;
      run += c;
    }
    s_cnt[depth] = run;
    ts->nlev = depth;
    ts->levels = depth;
  }
  __syncthreads();
  for (int j = tid; j < n; j += B) s_order[atomicAdd(&s_cur[s_level[j]], 1)] = j;
  __syncthreads();
  // the levels in order, on the resident quads
  const HarvestSink hv{};
  for (int L = 0; L < depth; ++L) {
    const int lo = s_cnt[L], hi = s_cnt[L + 1];
    for (int t = lo + tid; t < hi; t += B)
      kinetics_cell<false, true>(s_order[t], cells, rules, field, light, nbmask, events, p, hv, sI + tid * 17,
                                 sE + tid * 13, sD + tid * 13, nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
  }
  // fate (lib.rs:220-277): division targets in the pre-birth map
  CellParams pd = p;
  pd.stream = div_stream(ts);
  for (int i = tid; i <= n; i += B) {
    if (i < n) {
      div_target_cell(i, cells, events, cell_at, s_target, s_flag, pd);
    } else {
      s_target[i] = -1;
      s_flag[i] = 0;
    }
  }
  __syncthreads();
  int32_t* brank = s_level;
  block_exscan(s_flag, brank, n, s_w);
  for (int i = tid; i < n; i += B) birth_cell(i, cells, rules, s_target, brank, bcells, brules, s_btarget, claim, pd);
  int32_t* dflag = s_order;
  for (int i = tid; i < n; i += B) dflag[i] = events[i] == MARL_EVENT_DEATH ? 1 : 0;
  __syncthreads();
  block_exscan(dflag, s_P, n, s_w);
  const int nb = brank[n], nd = s_P[n], m0 = n - nd;
  const int holes = (nd > 0 && m0 > 0) ? s_P[m0] : 0;
  // the first birth claiming a voxel wins; then the claims are cleared
  int32_t* win = s_flag;
  for (int r = tid; r <= nb; r += B) win[r] = (r < nb && claim[s_btarget[r]] == (uint32_t)r) ? 1 : 0;
  __syncthreads();
  for (int r = tid; r < nb; r += B) claim[s_btarget[r]] = 0xffffffffu;
  block_exscan(win, s_wrank, nb, s_w);
  const int nw = s_wrank[nb];
  // deaths: head holes filled from the tail (closed form)
  int32_t* hole_dst = s_btarget;
  int32_t* hole_src = s_target;
  if (holes > 0) {
    for (int s0 = tid; s0 < m0; s0 += B) hole_source(s0, n, m0, events, s_P, hole_dst, hole_src);
    __syncthreads();
    for (int t = tid; t < holes * kRecChunks; t += B)
      move_chunk(t, hole_dst, hole_src, cells, rules, cells, rules, nullptr);
  }
  __syncthreads();
  // births appended after the survivors, in event order
  if (nw > 0) {
    int32_t* slot = s_wrank;  // each entry turns into its birth's slot in place
    for (int r = tid; r < nb; r += B) slot[r] = win[r] ? m0 + s_wrank[r] : 0;
    __syncthreads();
    for (int t = tid; t < nb * kRecChunks; t += B) move_chunk(t, slot, nullptr, bcells, brules, cells, rules, win);
  }
  if (tid == 0) {
    ts->nb = nb;
    ts->nd = nd;
    ts->m0 = m0;
    ts->holes = holes;
    ts->nw = nw;
  }
}

#undef GS_LOOP

constexpr int kBudgetBlocks = 512;

// The 40 FP64 totals of the resident state (empty-voxel field, occupied-voxel
// field, cell internals; occupancy from ctx->occ) reduced into `dst` (40
// doubles), or into the budget buffer's final slot when dst is null.  Fill
// kernels clear the partials, so the sequence can run inside a tick.
int enqueue_budget(marl_ctx* ctx, double* dst);

inline unsigned grid1(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// host orchestration

template <typename T>
int dev_alloc(marl_ctx* ctx, T** p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  MARL_HIP(ctx, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return MARL_OK;
}

int scan_i32(marl_ctx* ctx, const int32_t* in, int32_t* out, int n) {
  size_t need = 0;
  MARL_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, n, ctx->stream));
  if (need > ctx->scan_tmp_bytes) {
    if (ctx->scan_tmp) (void)hipFree(ctx->scan_tmp);
    ctx->scan_tmp = nullptr;
    MARL_HIP(ctx, hipMalloc(&ctx->scan_tmp, need * 2));
    ctx->scan_tmp_bytes = need * 2;
  }
  size_t have = ctx->scan_tmp_bytes;
  MARL_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp, have, in, out, n, ctx->stream));
  return MARL_OK;
}

int fill_i32(marl_ctx* ctx, int32_t* a, int32_t v, size_t n) {
  if (n == 0) return MARL_OK;
  const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, ctx->stream, a, v, n);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

int d2h_sync(marl_ctx* ctx, void* host, const void* dev, size_t bytes) {
  MARL_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

// Every block writes every one of its 40 partial slots, so nothing needs
// clearing between reductions.
int enqueue_budget(marl_ctx* ctx, double* dst) {
  if (!ctx->budget) MARL_HIP(ctx, hipMalloc(&ctx->budget, sizeof(double) * (kBudgetBlocks * kBudget + kBudget)));
  hipLaunchKernelGGL(k_budget_field, dim3(kBudgetBlocks), dim3(256), 0, ctx->stream, ctx->field[ctx->cur], ctx->occ,
                     ctx->nvox, ctx->budget);
  hipLaunchKernelGGL(k_budget_cells, dim3(kBudgetBlocks), dim3(256), 0, ctx->stream, ctx->cells[ctx->cbuf],
                     (int)ctx->n_cells, ctx->budget);
  hipLaunchKernelGGL(k_budget_final, dim3(1), dim3(64), 0, ctx->stream, ctx->budget, kBudgetBlocks,
                     dst ? dst : ctx->budget + kBudgetBlocks * kBudget);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// Per-context scratch that is sized by voxel count (allocated once).
int ensure_voxel_buffers(marl_ctx* ctx) {
  if (!ctx->cell_at) {
    if (int rc = dev_alloc(ctx, &ctx->cell_at, ctx->nvox)) return rc;
    if (int rc = fill_i32(ctx, ctx->cell_at, -1, ctx->nvox)) return rc;
  }
  if (!ctx->claim) {
    if (int rc = dev_alloc(ctx, &ctx->claim, ctx->nvox)) return rc;
    if (int rc = fill_i32(ctx, reinterpret_cast<int32_t*>(ctx->claim), -1, ctx->nvox)) return rc;
  }
  if (!ctx->level_count) {
    if (int rc = dev_alloc(ctx, &ctx->level_count, 2 * (kMaxLevels + 1))) return rc;
  }
  if (!ctx->misc) {
    if (int rc = dev_alloc(ctx, &ctx->misc, 64)) return rc;
  }
  return MARL_OK;
}

}  // namespace

// Grows every per-cell buffer to hold `need` cells, preserving the live ones.
int marl_ensure_cell_capacity(marl_ctx* ctx, int64_t need) {
  if (need <= ctx->cap) return MARL_OK;
  int64_t cap = std::max<int64_t>(need, 4096);
  cap = std::min<int64_t>(std::max<int64_t>(cap, ctx->cap * 2), (int64_t)ctx->nvox * 2 + 16);
  cap = std::max<int64_t>(cap, need);
  marl_cell_state* nc[2] = {nullptr, nullptr};
  marl_ruleset* nr[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; ++b) {
    MARL_HIP(ctx, hipMalloc(&nc[b], cap * sizeof(marl_cell_state)));
    MARL_HIP(ctx, hipMalloc(&nr[b], cap * sizeof(marl_ruleset)));
  }
  if (ctx->n_cells > 0) {
    MARL_HIP(ctx, hipMemcpyAsync(nc[0], ctx->cells[ctx->cbuf], ctx->n_cells * sizeof(marl_cell_state),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    MARL_HIP(ctx, hipMemcpyAsync(nr[0], ctx->rules[ctx->cbuf], ctx->n_cells * sizeof(marl_ruleset),
                                 hipMemcpyDeviceToDevice, ctx->stream));
  }
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < 2; ++b) {
    if (ctx->cells[b]) (void)hipFree(ctx->cells[b]);
    if (ctx->rules[b]) (void)hipFree(ctx->rules[b]);
    ctx->cells[b] = nc[b];
    ctx->rules[b] = nr[b];
  }
  ctx->cbuf = 0;
  if (int rc = dev_alloc(ctx, &ctx->nbmask, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->level, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->order, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->events, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->flag_scan, 4 * (cap + 1))) return rc;   // target, bflag, brank, dflag/P
  if (int rc = dev_alloc(ctx, &ctx->flag_scan2, 3 * (cap + 1))) return rc;  // btarget, win, wrank
  if (int rc = dev_alloc(ctx, &ctx->preds, (size_t)cap * kMaxPreds)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->npred, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->birth_cells, cap)) return rc;
  if (int rc = dev_alloc(ctx, &ctx->birth_rules, cap)) return rc;
  ctx->cap = cap;
  return MARL_OK;
}

int marl_launch_occupancy_from_cells(marl_ctx* ctx) {
  if (int rc = ensure_voxel_buffers(ctx)) return rc;
  if (int rc = fill_i32(ctx, ctx->cell_at, -1, ctx->nvox)) return rc;
  if (int rc = fill_i32(ctx, reinterpret_cast<int32_t*>(ctx->occ), 0, ctx->occ_words)) return rc;
  const int n = (int)ctx->n_cells;
  if (n > 0) {
    const CellParams p = cell_params(ctx, &ctx->last_sim, 0);
    hipLaunchKernelGGL(k_scatter_cells, dim3(grid1(n)), dim3(256), 0, ctx->stream, ctx->cells[ctx->cbuf], n,
                       ctx->cell_at, ctx->occ, p);
    MARL_HIP(ctx, hipGetLastError());
  }
  return MARL_OK;
}

// Field layout for the plain cell pass.  The quads in place cost each cell
// about 0.33 ns more than the AoS view, since a neighbour's species then span
// three cache lines instead of one or two.  The view costs whole-field
// transposes (48 bytes each way per voxel, about 15 ps per voxel at
// 256x256x128).  In a single-column tick the last substep writes the view
// (about +6 ps per voxel on that substep) and only the transpose back remains,
// beside fate: the tick is as fast either way near one cell per 20 voxels
// (measured at 256x256x128: 200k cells 1.958 vs 1.948 ms with quads, 400k
// 2.146 vs 2.141, 550k 2.289 vs 2.356, 1M 2.682 vs 2.997).  An ensemble
// column transposes both ways, which moves the crossover to about one cell
// per 10 voxels.  A band around the crossover keeps a captured tick from
// switching layouts back and forth.
bool marl_choose_cell_layout(marl_ctx* ctx) {
  if (ctx->cell_layout == MARL_CELL_LAYOUT_AOS) return ctx->kin_quads = false;
  if (ctx->cell_layout == MARL_CELL_LAYOUT_QUADS) return ctx->kin_quads = true;
  const double r = 10.0 * (double)ctx->n_cells / (double)std::max<int64_t>(ctx->nvox, 1);
  const double to_view = ctx->layout_view ? 0.6 : 1.25, to_quads = ctx->layout_view ? 0.35 : 0.8;
  if (ctx->kin_quads && r > to_view) ctx->kin_quads = false;
  else if (!ctx->kin_quads && r < to_quads) ctx->kin_quads = true;
  return ctx->kin_quads;
}

// Whether the plain or learned pass of this tick runs on the AoS view (the
// choice run_levels makes; the population is the same until the pass).
static bool pass_uses_view(marl_ctx* ctx, const CellParams& p) {
  const HarvestSink hv = surrogate_harvest_sink(ctx, p);
  const bool plain = ctx->kin_mode == MARL_KINETICS_PLAIN || !ctx->surr;
  return !(plain && !hv.count && marl_choose_cell_layout(ctx));
}

int marl_set_cell_layout(marl_ctx* ctx, int32_t layout) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (layout != MARL_CELL_LAYOUT_AUTO && layout != MARL_CELL_LAYOUT_AOS && layout != MARL_CELL_LAYOUT_QUADS)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "unknown cell-pass layout");
  ctx->cell_layout = layout;
  return MARL_OK;
}

namespace {

// The ordered cell pass (lib.rs:204-217) on the resident field, light and
// occupancy.  Leaves events in ctx->events; *levels_out = schedule depth.
// One level of the plain pass with the context's kinetics block size.
template <bool QUADS>
void launch_plain_level_q(marl_ctx* ctx, const int32_t* ord, int cnt, marl_cell_state* cells,
                          const marl_ruleset* rules, float* field, const CellParams& p) {
  const HarvestSink hv{};
  switch (ctx->kin_block) {
    case 64:
      hipLaunchKernelGGL((k_kinetics<false, QUADS, 64>), dim3(grid1(cnt, 64)), dim3(64), 0, ctx->stream, ord, cnt,
                         cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p, hv);
      break;
    case 256:
      hipLaunchKernelGGL((k_kinetics<false, QUADS, 256>), dim3(grid1(cnt, 256)), dim3(256), 0, ctx->stream, ord,
                         cnt, cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p, hv);
      break;
    default:
      hipLaunchKernelGGL((k_kinetics<false, QUADS, 128>), dim3(grid1(cnt, 128)), dim3(128), 0, ctx->stream, ord,
                         cnt, cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p, hv);
  }
}

void launch_plain_level(marl_ctx* ctx, bool quads, const int32_t* ord, int cnt, marl_cell_state* cells,
                        const marl_ruleset* rules, float* field, const CellParams& p) {
  if (quads) launch_plain_level_q<true>(ctx, ord, cnt, cells, rules, field, p);
  else launch_plain_level_q<false>(ctx, ord, cnt, cells, rules, field, p);
}

// The conflict schedule of the ordered pass (lib.rs:204-217): empty-neighbour
// masks, conflict levels and the cells sorted by level.  Reads only the cells,
// the occupancy and the voxel map, so it may run beside the field phase.
// On return off[L]..off[L+1] index level L's cells in ctx->order (off has
// depth + 1 entries; {0} when there are no cells).
int run_schedule(marl_ctx* ctx, const CellParams& p, std::vector<int32_t>& off) {
  const int n = (int)ctx->n_cells;
  off.assign(1, 0);
  if (n == 0) return MARL_OK;
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  hipLaunchKernelGGL(k_cell_prep, dim3(grid1(n)), dim3(256), 0, ctx->stream, cells, n, ctx->cell_at, ctx->occ,
                     ctx->nbmask, ctx->preds, (size_t)ctx->cap, ctx->npred, ctx->level, p);
  MARL_HIP(ctx, hipGetLastError());
  // Relaxation sweeps, each with its own change flag, in batches with one host
  // check each.  The first batch is as long as the previous tick needed (the
  // sweep that changed nothing included), so a schedule of steady depth costs
  // one batch and one check.  Extra sweeps after the fixed point change nothing.
  int32_t* changed = ctx->misc;
  int32_t* h = ctx->pinned;
  constexpr int kBatchMax = 32;
  int done = 0, batch = std::min(std::max(ctx->relax_sweeps, 2), kBatchMax);
  for (;;) {
    if (done > kMaxLevels + 2) return marl_fail(ctx, MARL_ERR_TRANSFER, "conflict schedule did not converge");
    if (int rc = fill_i32(ctx, changed, 0, batch)) return rc;
    for (int k = 0; k < batch; ++k)
      hipLaunchKernelGGL(k_level_relax, dim3(grid1(n)), dim3(256), 0, ctx->stream, done + k, n, ctx->preds,
                         (size_t)ctx->cap, ctx->npred, ctx->level, changed + k);
    MARL_HIP(ctx, hipGetLastError());
    if (int rc = d2h_sync(ctx, h, changed, sizeof(int32_t) * batch)) return rc;
    int quiet = -1;
    for (int k = 0; k < batch && quiet < 0; ++k)
      if (h[k] == 0) quiet = k;
    if (quiet >= 0) {
      ctx->relax_sweeps = done + quiet + 1;
      break;
    }
    // deeper than last tick: a short batch first, then doubling
    const int next = done == 0 ? 4 : std::min(2 * batch, kBatchMax);
    done += batch;
    batch = next;
  }
  int32_t* hist = ctx->level_count;
  if (int rc = fill_i32(ctx, hist, 0, kMaxLevels + 1)) return rc;
  const unsigned lb = std::min<unsigned>(grid1(n), kLevelBlocks);
  hipLaunchKernelGGL(k_level_hist, dim3(lb), dim3(256), 0, ctx->stream, n, ctx->level, hist);
  MARL_HIP(ctx, hipGetLastError());
  std::vector<int32_t> hh(kMaxLevels + 1);
  if (int rc = d2h_sync(ctx, hh.data(), hist, sizeof(int32_t) * (kMaxLevels + 1))) return rc;
  const int nlev = hh[0] + 1;
  if (nlev > kMaxLevels) return marl_fail(ctx, MARL_ERR_TRANSFER, "cell-pass schedule deeper than 4096 levels");
  off.assign(nlev + 1, 0);
  for (int L = 0; L < nlev; ++L) off[L + 1] = off[L] + hh[1 + L];
  int32_t* cursor = ctx->level_count + (kMaxLevels + 1);
  MARL_HIP(ctx, hipMemcpyAsync(cursor, off.data(), sizeof(int32_t) * nlev, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_level_scatter, dim3(lb), dim3(256), 0, ctx->stream, n, ctx->level, cursor, ctx->order);
  MARL_HIP(ctx, hipGetLastError());
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the copy reads `off` (pageable, staged at once) and `hh`
  return MARL_OK;
}

// The per-level launches of the ordered pass, after the schedule.  Leaves the
// events in ctx->events.  The pass works on an AoS view of the field (the
// 48-byte voxel records keep a neighbour's 12 species in one or two cache
// lines; the resident layout is species quads, best for the stencil), or on
// the quads in place when the population is sparse.  The surrogate and the
// pair harvest use the AoS view.  With back_done, the view is written back to
// the quads on ctx->aux after the levels, between back_start and back_done
// recorded there (the caller joins back_done before the field is next read);
// otherwise on ctx->stream.
int run_levels(marl_ctx* ctx, const CellParams& p, const std::vector<int32_t>& off, hipEvent_t back_start,
               hipEvent_t back_done) {
  const int nlev = (int)off.size() - 1;
  // the field phase may have left the new field in the view already
  const bool fresh = ctx->view_fresh;
  ctx->view_fresh = false;
  const HarvestSink hv = surrogate_harvest_sink(ctx, p);
  const bool plain = ctx->kin_mode == MARL_KINETICS_PLAIN || !ctx->surr;
  const bool quads = plain && !hv.count && marl_choose_cell_layout(ctx);
  if (fresh && quads) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell pass on the quads after a field phase into the view");
  if (nlev <= 0 && !fresh) {
    if (back_done) {
      MARL_HIP(ctx, hipEventRecord(back_start, ctx->aux));
      MARL_HIP(ctx, hipEventRecord(back_done, ctx->aux));
    }
    return MARL_OK;
  }
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  const marl_ruleset* rules = ctx->rules[ctx->cbuf];
  float* field = ctx->field[ctx->cur];
  if (!quads) {
    if (int rc = marl_ensure_staging(ctx)) return rc;
    if (!fresh)
      if (int rc = marl_transpose(ctx, false, ctx->field[ctx->cur])) return rc;
    field = ctx->staging;
  }
  for (int L = 0; L < nlev; ++L) {
    const int cnt = off[L + 1] - off[L];
    if (cnt == 0) continue;
    const dim3 grid(grid1(cnt, kKinBlock)), block(kKinBlock);
    const int32_t* ord = ctx->order + off[L];
    if (!plain) {
      if (int rc = surrogate_level(ctx, p, ord, cnt, cells, rules, field)) return rc;
    } else if (hv.count) {
      hipLaunchKernelGGL((k_kinetics<true, false>), grid, block, 0, ctx->stream, ord, cnt, cells, rules, field,
                         ctx->light, ctx->nbmask, ctx->events, p, hv);
    } else {
      launch_plain_level(ctx, quads, ord, cnt, cells, rules, field, p);
    }
  }
  MARL_HIP(ctx, hipGetLastError());
  if (back_done) {
    MARL_HIP(ctx, hipEventRecord(ctx->xev[0], ctx->stream));
    MARL_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->xev[0], 0));
    MARL_HIP(ctx, hipEventRecord(back_start, ctx->aux));
    if (!quads) {
      OnStream on(ctx, ctx->aux);
      if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) return rc;
    }
    MARL_HIP(ctx, hipEventRecord(back_done, ctx->aux));
  } else if (!quads) {
    if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) return rc;
  }
  return MARL_OK;
}

// The ordered cell pass (lib.rs:204-217) on the resident field, light and
// occupancy, on ctx->stream.  Leaves events in ctx->events; *levels_out =
// schedule depth; ev_sched (if given) is recorded after the schedule.
int run_cell_pass(marl_ctx* ctx, const CellParams& p, int32_t* levels_out, hipEvent_t ev_sched) {
  if (levels_out) *levels_out = 0;
  std::vector<int32_t> off;
  if (int rc = run_schedule(ctx, p, off)) return rc;
  if (ev_sched) MARL_HIP(ctx, hipEventRecord(ev_sched, ctx->stream));
  if (int rc = run_levels(ctx, p, off, nullptr, nullptr)) return rc;
  if (levels_out) *levels_out = (int)off.size() - 1;
  return MARL_OK;
}

// Fate, births with mutation, deaths and compaction (lib.rs:220-277).
int run_fate(marl_ctx* ctx, const CellParams& p, marl_tick_stats* st) {
  const int n = (int)ctx->n_cells;
  if (n == 0) return MARL_OK;
  const int64_t cap = ctx->cap;
  int32_t* target = ctx->flag_scan;
  int32_t* bflag = target + (cap + 1);
  int32_t* brank = bflag + (cap + 1);
  int32_t* P = brank + (cap + 1);
  int32_t* btarget = ctx->flag_scan2;
  int32_t* win = btarget + (cap + 1);
  int32_t* wrank = win + (cap + 1);
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  marl_ruleset* rules = ctx->rules[ctx->cbuf];

  hipLaunchKernelGGL(k_div_target, dim3(grid1(n)), dim3(256), 0, ctx->stream, n, cells, ctx->events, ctx->cell_at,
                     target, bflag, p);
  MARL_HIP(ctx, hipGetLastError());
  if (int rc = fill_i32(ctx, bflag + n, 0, 1)) return rc;
  if (int rc = scan_i32(ctx, bflag, brank, n + 1)) return rc;  // brank[n] = number of births
  hipLaunchKernelGGL(k_births, dim3((unsigned)std::min(n, kBirthGrid)), dim3(64), 0, ctx->stream, n, cells,
                     rules, target, brank, ctx->birth_cells, ctx->birth_rules, btarget, ctx->claim, p);
  MARL_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_death_flags, dim3(grid1(n)), dim3(256), 0, ctx->stream, n, ctx->events, win);
  MARL_HIP(ctx, hipGetLastError());
  if (int rc = fill_i32(ctx, win + n, 0, 1)) return rc;
  if (int rc = scan_i32(ctx, win, P, n + 1)) return rc;  // P[n] = deaths
  int32_t* h = ctx->pinned;
  MARL_HIP(ctx, hipMemcpyAsync(h, brank + n, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipMemcpyAsync(h + 1, P + n, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int nb = h[0], nd = h[1];
  int nw = 0;
  if (nb > 0) {
    hipLaunchKernelGGL(k_birth_winners, dim3(grid1(nb)), dim3(256), 0, ctx->stream, nb, btarget, ctx->claim, win);
    MARL_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_claim_reset, dim3(grid1(nb)), dim3(256), 0, ctx->stream, nb, btarget, ctx->claim);
    MARL_HIP(ctx, hipGetLastError());
    if (int rc = fill_i32(ctx, win + nb, 0, 1)) return rc;
    if (int rc = scan_i32(ctx, win, wrank, nb + 1)) return rc;
    if (int rc = d2h_sync(ctx, h + 2, wrank + nb, sizeof(int32_t))) return rc;
    nw = h[2];
  }
  const int m0 = n - nd;
  const int64_t n_new = (int64_t)m0 + nw;
  // deaths: fill the head holes in place from the tail (lib.rs:257-267)
  if (nd > 0 && m0 > 0) {
    // scratch reuse: per-cell targets and per-birth targets are consumed by now
    int32_t* hole_dst = btarget;
    int32_t* hole_src = target;
    // hole count = deaths below m0 = P[m0]
    if (int rc = d2h_sync(ctx, h + 3, P + m0, sizeof(int32_t))) return rc;
    const int holes = h[3];
    if (holes > 0) {
      hipLaunchKernelGGL(k_hole_sources, dim3(grid1(m0)), dim3(256), 0, ctx->stream, n, m0, ctx->events, P,
                         hole_dst, hole_src);
      MARL_HIP(ctx, hipGetLastError());
      hipLaunchKernelGGL(k_move_records, dim3(grid1((int64_t)holes * kRecChunks)), dim3(256), 0, ctx->stream,
                         holes, hole_dst, hole_src, cells, rules, cells, rules, nullptr);
      MARL_HIP(ctx, hipGetLastError());
    }
  }
  // births appended in event order after the survivors (lib.rs:270-277)
  if (nb > 0 && nw > 0) {
    int32_t* slot = bflag;  // birth flags were consumed by the rank scan
    hipLaunchKernelGGL(k_birth_slots, dim3(grid1(nb)), dim3(256), 0, ctx->stream, nb, m0, win, wrank, slot);
    MARL_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_move_records, dim3(grid1((int64_t)nb * kRecChunks)), dim3(256), 0, ctx->stream, nb, slot,
                       nullptr, ctx->birth_cells, ctx->birth_rules, cells, rules, win);
    MARL_HIP(ctx, hipGetLastError());
  }
  ctx->n_cells = n_new;
  if (st) {
    st->divisions = nb;
    st->deaths = nd;
    st->births_added = nw;
    st->n_cells = n_new;
  }
  return MARL_OK;
}

// Reserves hipcub scan storage for n items (no allocation may happen while a
// graph is being captured).
int scan_reserve(marl_ctx* ctx, int n) {
  size_t need = 0;
  MARL_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, (const int32_t*)nullptr, (int32_t*)nullptr, n,
                                                 ctx->stream));
  if (need > ctx->scan_tmp_bytes) {
    if (ctx->scan_tmp) (void)hipFree(ctx->scan_tmp);
    ctx->scan_tmp = nullptr;
    MARL_HIP(ctx, hipMalloc(&ctx->scan_tmp, need * 2));
    ctx->scan_tmp_bytes = need * 2;
  }
  return MARL_OK;
}

inline unsigned gs_grid(int64_t n, int block = 256, int cap_blocks = 2048) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, cap_blocks));
}

}  // namespace

// The device-driven tick in three parts, each enqueued on ctx->stream and
// capturable (no host synchronisation, no allocation, no memset nodes):
// counters and occupancy; the field phase; the cell phase.
int marl_enqueue_tick_begin(marl_ctx* ctx, const marl_sim_params* sim) {
  TickDev* ts = static_cast<TickDev*>(ctx->tdev);
  const CellParams p = cell_params(ctx, sim, 0);
  hipLaunchKernelGGL(k_g_begin, dim3(1), dim3(64), 0, ctx->stream, ts, ctx->cap);
  // occupancy from the live cells (lib.rs:174-179)
  if (int rc = fill_i32(ctx, ctx->cell_at, -1, ctx->nvox)) return rc;
  if (int rc = fill_i32(ctx, reinterpret_cast<int32_t*>(ctx->occ), 0, ctx->occ_words)) return rc;
  hipLaunchKernelGGL(k_g_scatter_cells, dim3(gs_grid(ctx->cap)), dim3(256), 0, ctx->stream, ctx->cells[ctx->cbuf],
                     ts, ctx->cell_at, ctx->occ, p);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// Substeps with the boundary sources fused into the first, then light (fused
// into the last substep where the tile allows).
int marl_enqueue_field_phase(marl_ctx* ctx, const marl_sim_params* sim) {
  return marl_launch_field_phase(ctx, sim, true);
}

// Schedule, the ordered pass, fate, births, deaths and compaction: one
// workgroup for a small population (ctx->graph_small), else the multi-launch
// sequence.
int marl_enqueue_cell_phase(marl_ctx* ctx, const marl_sim_params* sim) {
  if (ctx->graph_small) return marl_enqueue_cell_phase_small(ctx, sim);
  if (int rc = marl_enqueue_schedule(ctx, sim)) return rc;
  return marl_enqueue_pass_and_fate(ctx, sim, false);
}

// The one-workgroup cell phase and the counters' finish, on ctx->stream
// (TickDev staged; the resident quads hold the field).
int marl_enqueue_cell_phase_small(marl_ctx* ctx, const marl_sim_params* sim) {
  if (ctx->view_fresh) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell pass on the quads after a field phase into the view");
  TickDev* ts = static_cast<TickDev*>(ctx->tdev);
  const CellParams p = cell_params(ctx, sim, 0);  // the RNG stream comes from ts->tick
  hipLaunchKernelGGL(k_g_cell_phase_small, dim3(1), dim3(kSmallThreads), 0, ctx->stream, ts, ctx->cells[ctx->cbuf],
                     ctx->rules[ctx->cbuf], ctx->field[ctx->cur], ctx->light, ctx->cell_at, ctx->occ, ctx->nbmask,
                     ctx->preds, (size_t)ctx->cap, ctx->npred, ctx->events, ctx->birth_cells, ctx->birth_rules,
                     ctx->claim, p);
  hipLaunchKernelGGL(k_g_finish, dim3(1), dim3(1), 0, ctx->stream, ts);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// The conflict schedule (masks, levels, level lists) on ctx->stream.
int marl_enqueue_schedule(marl_ctx* ctx, const marl_sim_params* sim) {
  TickDev* ts = static_cast<TickDev*>(ctx->tdev);
  const int64_t cap = ctx->cap;
  const CellParams p = cell_params(ctx, sim, 0);
  hipStream_t st = ctx->stream;
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  const unsigned gc = gs_grid(cap);
  hipLaunchKernelGGL(k_g_cell_prep, dim3(gc), dim3(256), 0, st, cells, ts, ctx->cell_at, ctx->occ, ctx->nbmask,
                     ctx->preds, (size_t)cap, ctx->npred, ctx->level, p);
  for (int k = 0; k < ctx->graph_sweeps; ++k)
    hipLaunchKernelGGL(k_g_level_relax, dim3(gc), dim3(256), 0, st, k, ts, ctx->preds, (size_t)cap, ctx->npred,
                       ctx->level);
  hipLaunchKernelGGL(k_g_relax_check, dim3(1), dim3(1), 0, st, ts, ctx->graph_sweeps);
  int32_t* hist = ctx->level_count;
  int32_t* cursor = ctx->level_count + (kMaxLevels + 1);
  if (int rc = fill_i32(ctx, hist, 0, kMaxLevels + 1)) return rc;
  hipLaunchKernelGGL(k_g_level_hist, dim3(std::min<unsigned>(gc, kLevelBlocks)), dim3(256), 0, st, ts, ctx->level,
                     hist);
  hipLaunchKernelGGL(k_g_level_offsets, dim3(1), dim3(1), 0, st, ts, hist, cursor, ctx->graph_levels);
  hipLaunchKernelGGL(k_g_level_scatter, dim3(std::min<unsigned>(gc, kLevelBlocks)), dim3(256), 0, st, ts,
                     ctx->level, cursor, ctx->order);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// The ordered pass, fate, births, deaths and compaction on ctx->stream, after
// the schedule and the field phase.  With fork_back the AoS view goes back to
// the quads on ctx->aux while fate runs (fate reads no field), joined on
// ctx->stream before the part ends.
int marl_enqueue_pass_and_fate(marl_ctx* ctx, const marl_sim_params* sim, bool fork_back) {
  TickDev* ts = static_cast<TickDev*>(ctx->tdev);
  const int64_t cap = ctx->cap;
  const CellParams p = cell_params(ctx, sim, 0);  // the RNG stream comes from ts->tick
  hipStream_t st = ctx->stream;
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  marl_ruleset* rules = ctx->rules[ctx->cbuf];
  const int kb = ctx->kin_block;
  const unsigned gc = gs_grid(cap), gk = gs_grid(cap, kb, 16384 * 64 / kb), gf = gs_grid(cap + 1);
  // the ordered pass on the AoS view, or on the quads in place (the layout
  // chosen when the tick was prepared; it is part of the graph's key)
  auto level = [&](int L, float* field, bool quads) {
    if (quads) {
      if (kb == 64) hipLaunchKernelGGL((k_g_kinetics<true, 64>), dim3(gk), dim3(64), 0, st, L, ts, ctx->order,
                                       cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p);
      else if (kb == 256) hipLaunchKernelGGL((k_g_kinetics<true, 256>), dim3(gk), dim3(256), 0, st, L, ts,
                                             ctx->order, cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p);
      else hipLaunchKernelGGL((k_g_kinetics<true, 128>), dim3(gk), dim3(128), 0, st, L, ts, ctx->order, cells,
                              rules, field, ctx->light, ctx->nbmask, ctx->events, p);
    } else {
      if (kb == 64) hipLaunchKernelGGL((k_g_kinetics<false, 64>), dim3(gk), dim3(64), 0, st, L, ts, ctx->order,
                                       cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p);
      else if (kb == 256) hipLaunchKernelGGL((k_g_kinetics<false, 256>), dim3(gk), dim3(256), 0, st, L, ts,
                                             ctx->order, cells, rules, field, ctx->light, ctx->nbmask, ctx->events, p);
      else hipLaunchKernelGGL((k_g_kinetics<false, 128>), dim3(gk), dim3(128), 0, st, L, ts, ctx->order, cells,
                              rules, field, ctx->light, ctx->nbmask, ctx->events, p);
    }
  };
  const bool fresh = ctx->view_fresh;  // the field phase wrote the view
  ctx->view_fresh = false;
  if (ctx->kin_quads) {
    if (fresh) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell pass on the quads after a field phase into the view");
    for (int L = 0; L < ctx->graph_levels; ++L) level(L, ctx->field[ctx->cur], true);
  } else {
    if (!fresh)
      if (int rc = marl_transpose(ctx, false, ctx->field[ctx->cur])) return rc;
    for (int L = 0; L < ctx->graph_levels; ++L) level(L, ctx->staging, false);
    if (fork_back) {
      MARL_HIP(ctx, hipEventRecord(ctx->xev[2], st));
      MARL_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->xev[2], 0));
      OnStream on(ctx, ctx->aux);
      if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) return rc;
      MARL_HIP(ctx, hipEventRecord(ctx->xev[3], ctx->aux));
    } else if (int rc = marl_transpose(ctx, true, ctx->field[ctx->cur])) {
      return rc;
    }
  }
  MARL_HIP(ctx, hipGetLastError());
  // fate, births with mutation, deaths and compaction
  int32_t* target = ctx->flag_scan;
  int32_t* bflag = target + (cap + 1);
  int32_t* brank = bflag + (cap + 1);
  int32_t* P = brank + (cap + 1);
  int32_t* btarget = ctx->flag_scan2;
  int32_t* win = btarget + (cap + 1);
  int32_t* wrank = win + (cap + 1);
  hipLaunchKernelGGL(k_g_div_target, dim3(gf), dim3(256), 0, st, ts, cap, cells, ctx->events, ctx->cell_at, target,
                     bflag, p);
  if (int rc = scan_i32(ctx, bflag, brank, (int)cap + 1)) return rc;
  hipLaunchKernelGGL(k_g_births, dim3((unsigned)std::min<int64_t>(std::max<int64_t>(cap, 1), kBirthGrid)), dim3(64), 0, st,
                     ts, cells, rules, target, brank, ctx->birth_cells, ctx->birth_rules, btarget, ctx->claim, p);
  hipLaunchKernelGGL(k_g_death_flags, dim3(gf), dim3(256), 0, st, ts, cap, ctx->events, win);
  if (int rc = scan_i32(ctx, win, P, (int)cap + 1)) return rc;
  hipLaunchKernelGGL(k_g_counts, dim3(1), dim3(1), 0, st, ts, brank, P);
  hipLaunchKernelGGL(k_g_birth_winners, dim3(gf), dim3(256), 0, st, ts, cap, btarget, ctx->claim, win);
  hipLaunchKernelGGL(k_g_claim_reset, dim3(gc), dim3(256), 0, st, ts, btarget, ctx->claim);
  if (int rc = scan_i32(ctx, win, wrank, (int)cap + 1)) return rc;
  int32_t* hole_dst = btarget;
  int32_t* hole_src = target;
  hipLaunchKernelGGL(k_g_hole_sources, dim3(gc), dim3(256), 0, st, ts, ctx->events, P, wrank, hole_dst, hole_src);
  hipLaunchKernelGGL(k_g_move_holes, dim3(gs_grid(cap * kRecChunks, 256, 8192)), dim3(256), 0, st, ts, hole_dst,
                     hole_src, cells, rules);
  int32_t* slot = bflag;
  hipLaunchKernelGGL(k_g_birth_slots, dim3(gc), dim3(256), 0, st, ts, win, wrank, slot);
  hipLaunchKernelGGL(k_g_move_births, dim3(gs_grid(cap * kRecChunks, 256, 8192)), dim3(256), 0, st, ts, slot,
                     ctx->birth_cells, ctx->birth_rules, cells, rules, win);
  hipLaunchKernelGGL(k_g_finish, dim3(1), dim3(1), 0, st, ts);
  MARL_HIP(ctx, hipGetLastError());
  if (fork_back && !ctx->kin_quads) MARL_HIP(ctx, hipStreamWaitEvent(st, ctx->xev[3], 0));
  return MARL_OK;
}

// What a captured tick depends on: parameters, tile, buffers and their parity.
std::vector<uint8_t> marl_graph_key(const marl_ctx* ctx, const marl_sim_params* sim) {
  struct K {
    marl_sim_params sim;
    int64_t cap;
    const void *cells, *rules, *f0, *f1, *staging;
    int32_t tile, cur, cbuf, sweeps, levels, kin_quads, kin_block, overlap, small;
  } k;
  memset(&k, 0, sizeof(k));
  k.sim = *sim;
  k.cap = ctx->cap;
  k.cells = ctx->cells[ctx->cbuf];
  k.rules = ctx->rules[ctx->cbuf];
  k.f0 = ctx->field[0];
  k.f1 = ctx->field[1];
  k.staging = ctx->staging;
  k.tile = ctx->tile;
  k.cur = ctx->cur;
  k.cbuf = ctx->cbuf;
  k.sweeps = ctx->graph_small ? 0 : ctx->graph_sweeps;  // the one-workgroup phase has no budget
  k.levels = ctx->graph_small ? 0 : ctx->graph_levels;
  k.kin_quads = ctx->kin_quads ? 1 : 0;
  k.kin_block = ctx->kin_block;
  k.overlap = ctx->graph_overlap ? 1 : 0;
  k.small = ctx->graph_small ? 1 : 0;
  const uint8_t* b = reinterpret_cast<const uint8_t*>(&k);
  return std::vector<uint8_t>(b, b + sizeof(k));
}

// Buffers, capacity for this tick's births, scan storage and the device
// counters of a device-driven tick (allocations happen here, never during
// capture).
int marl_graph_prepare(marl_ctx* ctx, const marl_sim_params* sim) {
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (ctx->kin_mode != MARL_KINETICS_PLAIN || (ctx->surr && ctx->surr->harvesting))
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "the captured tick needs plain kinetics and no harvest");
  ctx->last_sim = *sim;
  if (int rc = ensure_voxel_buffers(ctx)) return rc;
  if (int rc = marl_ensure_staging(ctx)) return rc;
  if (!ctx->tdev) MARL_HIP(ctx, hipMalloc(&ctx->tdev, sizeof(TickDev)));
  const int64_t n = ctx->n_cells;
  marl_choose_cell_layout(ctx);
  // the one-workgroup cell phase for small populations (it works on the
  // resident quads), with hysteresis so that a population near the bound is
  // not captured again every tick; a tick that starts above kSmallCells is
  // finished by the host
  if (ctx->graph_small && (n > kSmallCells * 3 / 4 || !ctx->kin_quads)) ctx->graph_small = false;
  else if (!ctx->graph_small && n <= kSmallCells / 2 && ctx->kin_quads) ctx->graph_small = true;
  // room for this tick's births; generous so that growth rarely re-captures
  // (in small mode for any population up to kSmallCells, so that a run of
  // ticks never outgrows it)
  int64_t room = std::max<int64_t>(2 * n + 1024, 4 * n);
  if (ctx->graph_small) room = std::max<int64_t>(room, 2 * kSmallCells + 1024);
  if (int rc = marl_ensure_cell_capacity(ctx, room)) return rc;
  return scan_reserve(ctx, (int)ctx->cap + 1);
}

static TickDev* host_tickdev(marl_ctx* ctx) {
  static_assert(sizeof(TickDev) <= 2048, "TickDev fits the pinned scratch");
  return reinterpret_cast<TickDev*>(reinterpret_cast<uint8_t*>(ctx->pinned) + 2048);
}

// Before a replay: the live count and tick index into the device counters,
// as kernel arguments (k_stage_batch with one column; the host copy of the
// counters is kept in step for the host-finish path).
int marl_graph_stage(marl_ctx* ctx, hipStream_t st, int64_t tick) {
  marl_ctx* one[1] = {ctx};
  return marl_graph_stage_many(one, 1, st, tick);
}

// After a replay: the counters back (the caller synchronises st).
int marl_graph_fetch(marl_ctx* ctx, hipStream_t st) {
  MARL_HIP(ctx, hipMemcpyAsync(host_tickdev(ctx), ctx->tdev, sizeof(TickDev), hipMemcpyDeviceToHost, st));
  return MARL_OK;
}

namespace {

// Batched staging: (counters, live count, tick) of up to 64 columns by value.
struct StageBatch {
  TickDev* ts[64];
  int32_t n[64];
  uint64_t tick;
  int32_t count;
};

__global__ void k_stage_batch(StageBatch b) {
  const int i = threadIdx.x;
  if (i < b.count) {
    b.ts[i]->n = b.n[i];
    b.ts[i]->tick = b.tick;
  }
}

// One block per column: its counters into slot c of a contiguous array.
__global__ void k_gather_tickdevs(TickDev* const* __restrict__ src, TickDev* __restrict__ out) {
  constexpr int kWords = sizeof(TickDev) / 4;
  const int32_t* s = reinterpret_cast<const int32_t*>(src[blockIdx.x]);
  int32_t* d = reinterpret_cast<int32_t*>(out + blockIdx.x);
  for (int k = threadIdx.x; k < kWords; k += blockDim.x) d[k] = s[k];
}
static_assert(sizeof(TickDev) % 4 == 0, "TickDev words");

}  // namespace

size_t marl_tickdev_bytes() { return sizeof(TickDev); }

int marl_graph_stage_many(marl_ctx* const* cols, int n, hipStream_t st, int64_t tick) {
  for (int c0 = 0; c0 < n; c0 += 64) {
    StageBatch b{};
    b.count = std::min(64, n - c0);
    b.tick = (uint64_t)tick;
    for (int i = 0; i < b.count; ++i) {
      marl_ctx* ctx = cols[c0 + i];
      b.ts[i] = static_cast<TickDev*>(ctx->tdev);
      b.n[i] = (int32_t)ctx->n_cells;
      TickDev* hts = host_tickdev(ctx);  // kept in step with the device for the host-finish path
      hts->n = b.n[i];
      hts->tick = b.tick;
    }
    hipLaunchKernelGGL(k_stage_batch, dim3(1), dim3(64), 0, st, b);
    MARL_HIP(cols[c0], hipGetLastError());
  }
  return MARL_OK;
}

int marl_graph_fetch_many(marl_ctx* const* cols, int n, hipStream_t st, void* const* dev_srcs, void* dev_out,
                          void* host_out) {
  if (n <= 0) return MARL_OK;
  hipLaunchKernelGGL(k_gather_tickdevs, dim3(n), dim3(64), 0, st, reinterpret_cast<TickDev* const*>(dev_srcs),
                     static_cast<TickDev*>(dev_out));
  MARL_HIP(cols[0], hipGetLastError());
  MARL_HIP(cols[0], hipMemcpyAsync(host_out, dev_out, sizeof(TickDev) * (size_t)n, hipMemcpyDeviceToHost, st));
  return MARL_OK;
}

void marl_graph_fetch_finish_many(marl_ctx* const* cols, int n, const void* host_out) {
  const TickDev* h = static_cast<const TickDev*>(host_out);
  for (int c = 0; c < n; ++c) memcpy(host_tickdev(cols[c]), h + c, sizeof(TickDev));
}

// After the fetch completed: the tick's outcome, finishing the cell phase on
// the host-driven path if the device skipped it.
// Automatic schedule budget of the captured tick: follow the sweeps and levels
// the schedule needed, with slack, so that the graph launches few idle sweep
// and level kernels (each costs a launch even when it exits at once).  The
// slack is need / 2 + 2, between 3 and 8.  A budget within 1 of the need
// grows at once to the need plus twice the slack (a growing population
// deepens its schedule by a level or two per tick, and a tick over budget is
// finished by the host); a budget more than three slacks above the need
// shrinks to the need plus one slack, so that the graph is rarely captured
// again.
static void adapt_graph_limits(marl_ctx* ctx, int need_sweeps, int need_levels) {
  if (!ctx->graph_auto) return;
  auto fit = [](int cur, int need, int cap) {
    const int slack = std::min(8, std::max(3, need / 2 + 2));
    if (need + 1 >= cur) return std::min(cap, need + 2 * slack);
    return cur > need + 3 * slack ? std::min(cap, need + slack) : cur;
  };
  ctx->graph_sweeps = fit(ctx->graph_sweeps, std::max(need_sweeps, 1), kGraphSweeps);
  ctx->graph_levels = fit(ctx->graph_levels, std::max(need_levels, 1), kGraphLevels);
}

int marl_graph_post(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick, marl_tick_stats* st,
                    int32_t* host_finished) {
  const TickDev* hts = host_tickdev(ctx);
  *st = marl_tick_stats{};
  if (hts->err == 0) {
    ctx->n_cells = hts->n;
    st->divisions = hts->divisions;
    st->deaths = hts->deaths;
    st->births_added = hts->births_added;
    st->levels = hts->levels;
    int quiet = 0;  // the first sweep that changed nothing (the check found one)
    while (quiet < ctx->graph_sweeps - 1 && hts->chg[quiet] != 0) ++quiet;
    adapt_graph_limits(ctx, quiet + 1, hts->levels);
  } else {
    ++*host_finished;
    if (int rc = marl_ensure_cell_capacity(ctx, 2 * ctx->n_cells + 1024)) return rc;
    const CellParams p = cell_params(ctx, sim, (uint64_t)tick);
    int32_t levels = 0;
    if (int rc = run_cell_pass(ctx, p, &levels, nullptr)) return rc;
    if (int rc = run_fate(ctx, p, st)) return rc;
    MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    st->levels = levels;
    adapt_graph_limits(ctx, ctx->relax_sweeps, levels);
  }
  st->n_cells = ctx->n_cells;
  return MARL_OK;
}

namespace {

int drop_graph(marl_ctx* ctx, int slot) {
  if (ctx->gexec[slot]) (void)hipGraphExecDestroy(ctx->gexec[slot]);
  if (ctx->graph[slot]) (void)hipGraphDestroy(ctx->graph[slot]);
  ctx->gexec[slot] = nullptr;
  ctx->graph[slot] = nullptr;
  ctx->graph_key[slot].clear();
  return MARL_OK;
}

// Capture one device-driven tick for the current field parity.  Capture
// records work without running it, so the host-side parity is restored and
// the number of buffer swaps the graph performs is remembered.
int capture_tick(marl_ctx* ctx, const marl_sim_params* sim) {
  const int slot = ctx->cur;
  drop_graph(ctx, slot);
  MARL_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  // counters and occupancy; then the schedule on ctx->aux beside the field
  // phase (the schedule reads only cells and occupancy); the pass after both
  int rc = marl_enqueue_tick_begin(ctx, sim);
  const bool overlap = ctx->graph_overlap;
  bool forked = false;
  if (rc == MARL_OK && overlap) {
    if (hipEventRecord(ctx->xev[1], ctx->stream) == hipSuccess &&
        hipStreamWaitEvent(ctx->aux, ctx->xev[1], 0) == hipSuccess) {
      forked = true;
      OnStream on(ctx, ctx->aux);
      rc = marl_enqueue_schedule(ctx, sim);
    } else {
      rc = marl_fail(ctx, MARL_ERR_TRANSFER, "capture: fork of the schedule stream failed");
    }
  }
  // a pass on the AoS view takes the last substep's output directly
  if (rc == MARL_OK) rc = marl_launch_field_phase(ctx, sim, true, ctx->kin_quads ? nullptr : ctx->staging);
  const int cur_after = ctx->cur;
  if (forked) {  // join the schedule stream on every path, so that the capture can end
    const hipError_t j1 = hipEventRecord(ctx->xev[0], ctx->aux);
    const hipError_t j2 = j1 == hipSuccess ? hipStreamWaitEvent(ctx->stream, ctx->xev[0], 0) : j1;
    if (rc == MARL_OK && j2 != hipSuccess) rc = marl_hip_fail(ctx, j2, "capture: join of the schedule stream");
  }
  if (rc == MARL_OK) rc = overlap ? marl_enqueue_pass_and_fate(ctx, sim, true) : marl_enqueue_cell_phase(ctx, sim);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->cur = slot;
  ctx->view_fresh = false;  // nothing ran: the captured work starts from the quads
  if (rc != MARL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MARL_HIP(ctx, e);
  ctx->graph[slot] = g;
  MARL_HIP(ctx, hipGraphInstantiate(&ctx->gexec[slot], g, nullptr, nullptr, 0));
  ctx->graph_key[slot] = marl_graph_key(ctx, sim);
  ctx->graph_flip[slot] = cur_after != slot;
  return MARL_OK;
}

int check_cells_host(marl_ctx* ctx, const marl_cell_state* cells, size_t n) {
  std::vector<uint8_t> seen(ctx->nvox, 0);
  for (size_t i = 0; i < n; ++i) {
    const uint16_t* ps = cells[i].pos;
    if (ps[0] >= ctx->d.x || ps[1] >= ctx->d.y || ps[2] >= ctx->d.z)
      return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell position out of bounds");
    const size_t v = ((size_t)ps[2] * ctx->d.y + ps[1]) * ctx->d.x + ps[0];
    if (seen[v]) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "two cells share a voxel");
    seen[v] = 1;
  }
  return MARL_OK;
}

}  // namespace

// hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27), one pair per
// thread; the pairs are independent (checked on the host).
__global__ void k_hgt(const int32_t* __restrict__ pairs, int np, marl_ruleset* __restrict__ rules, uint64_t seed) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const marl_ruleset& d = rules[pairs[2 * k]];
  marl_ruleset& r = rules[pairs[2 * k + 1]];
  int active[MARL_R_MAX];
  int na = 0;
  for (int q = 0; q < MARL_R_MAX; ++q)
    if (fabsf(d.reactions[q].v_max) > 1e-9f) active[na++] = q;
  if (na == 0) return;
  MarlRng g;
  g.init(seed, MARL_TAG_HGT << 32, (uint64_t)k);
  const int src = active[g.range((uint32_t)na)];
  int tgt = -1;
  for (int q = 0; q < MARL_R_MAX && tgt < 0; ++q)
    if (fabsf(r.reactions[q].v_max) < 1e-9f) tgt = q;
  if (tgt < 0) tgt = (int)g.range(MARL_R_MAX);
  r.reactions[tgt] = d.reactions[src];
}

extern "C" {

int marl_upload_cells(marl_ctx* ctx, const marl_cell_state* cells, const marl_ruleset* rules, size_t n) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (n > 0 && (!cells || !rules)) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null cell buffers");
  if (n > ctx->nvox) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "more cells than voxels");
  if (int rc = check_cells_host(ctx, cells, n)) return rc;
  if (int rc = ensure_voxel_buffers(ctx)) return rc;
  ctx->n_cells = 0;
  if (int rc = marl_ensure_cell_capacity(ctx, std::max<int64_t>(2 * (int64_t)n, 1024))) return rc;
  if (n > 0) {
    MARL_HIP(ctx, hipMemcpyAsync(ctx->cells[ctx->cbuf], cells, n * sizeof(marl_cell_state), hipMemcpyHostToDevice,
                                 ctx->stream));
    MARL_HIP(ctx, hipMemcpyAsync(ctx->rules[ctx->cbuf], rules, n * sizeof(marl_ruleset), hipMemcpyHostToDevice,
                                 ctx->stream));
  }
  ctx->n_cells = (int64_t)n;
  if (int rc = marl_launch_occupancy_from_cells(ctx)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_cell_count(marl_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return MARL_ERR_INVALID_INPUT;
  *n = ctx->n_cells;
  return MARL_OK;
}

int marl_download_cells(marl_ctx* ctx, marl_cell_state* cells, marl_ruleset* rules, size_t capacity,
                        int64_t* n_out) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  const size_t n = (size_t)ctx->n_cells;
  if (n_out) *n_out = (int64_t)n;
  if (n == 0) return MARL_OK;
  if (capacity < n || !cells) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "download capacity below cell count");
  std::vector<marl_cell_state> tc(n);
  MARL_HIP(ctx, hipMemcpyAsync(tc.data(), ctx->cells[ctx->cbuf], n * sizeof(marl_cell_state), hipMemcpyDeviceToHost,
                               ctx->stream));
  std::vector<marl_ruleset> tr;
  if (rules) {
    tr.resize(n);
    MARL_HIP(ctx, hipMemcpyAsync(tr.data(), ctx->rules[ctx->cbuf], n * sizeof(marl_ruleset), hipMemcpyDeviceToHost,
                                 ctx->stream));
  }
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(cells, tc.data(), n * sizeof(marl_cell_state));
  if (rules) memcpy(rules, tr.data(), n * sizeof(marl_ruleset));
  return MARL_OK;
}

int marl_upload_light(marl_ctx* ctx, const float* light, size_t n) {
  if (!ctx || !light) return MARL_ERR_INVALID_INPUT;
  if (n != ctx->nvox) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "light length differs from voxel count");
  MARL_HIP(ctx, hipMemcpyAsync(ctx->light, light, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MARL_OK;
}

int marl_cell_pass(marl_ctx* ctx, const marl_sim_params* sim, uint8_t* events, size_t n_events, int32_t* levels) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  const size_t n = (size_t)ctx->n_cells;
  if (events && n_events < n) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "events buffer shorter than cell count");
  ctx->last_sim = *sim;
  if (int rc = marl_launch_occupancy_from_cells(ctx)) return rc;
  const CellParams p = cell_params(ctx, sim, 0);
  if (int rc = run_cell_pass(ctx, p, levels, nullptr)) return rc;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (events && n > 0) {
    std::vector<uint8_t> tmp(n);
    if (int rc = d2h_sync(ctx, tmp.data(), ctx->events, n)) return rc;
    memcpy(events, tmp.data(), n);
  }
  return MARL_OK;
}

int marl_tick(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick0, int32_t n_ticks, marl_tick_stats* stats) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  if (n_ticks < 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "negative tick count");
  if (int rc = marl_fault_tick(ctx)) return rc;
  ctx->last_sim = *sim;
  ctx->layout_view = true;
  if (int rc = ensure_voxel_buffers(ctx)) return rc;
  if (ctx->phase_budget && !ctx->phase_budget_dev)
    MARL_HIP(ctx, hipMalloc(&ctx->phase_budget_dev, sizeof(double) * 4 * kBudget));
  if (int rc = marl_ensure_aux(ctx)) return rc;
  constexpr int kEv = 9;
  hipEvent_t ev[kEv] = {};
  for (auto& e : ev) MARL_HIP(ctx, hipEventCreate(&e));
  struct Guard {
    hipEvent_t* e;
    ~Guard() {
      for (int k = 0; k < kEv; ++k)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } guard{ev};
  // every exit, error paths included, waits for the second stream, so that no
  // transpose of this call is still writing the field when the call returns
  // (after a normal tick it has finished already)
  struct AuxJoin {
    hipStream_t aux;
    ~AuxJoin() { (void)hipStreamSynchronize(aux); }
  } aux_join{ctx->aux};
  // a tick that fails between a field phase into the view and the pass's
  // transpose back restores the quads from the view (the field phase's result)
  struct ViewBack {
    marl_ctx* c;
    ~ViewBack() {
      if (!c->view_fresh) return;
      c->view_fresh = false;
      (void)marl_transpose(c, true, c->field[c->cur]);
      (void)hipStreamSynchronize(c->stream);
    }
  } view_back{ctx};
  marl_tick_stats st{};
  std::vector<int32_t> off;
  for (int t = 0; t < n_ticks; ++t) {
    st = marl_tick_stats{};
    if (int rc = marl_ensure_cell_capacity(ctx, 2 * ctx->n_cells + 1024)) return rc;
    const CellParams p = cell_params(ctx, sim, (uint64_t)(tick0 + t));
    // With the phase budget on, each phase finishes before the next begins (the
    // budget reads the field after the cell pass); otherwise the schedule runs
    // on ctx->aux beside the field phase and the transpose back beside fate.
    const bool overlap = !ctx->phase_budget;
    MARL_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
    // occupancy from the live cells (lib.rs:174-179); sources fused into the
    // first substep (field.rs:231-250 then :212-217); light (lib.rs:199),
    // fused into the last substep where the tile allows
    if (int rc = marl_launch_occupancy_from_cells(ctx)) return rc;
    if (ctx->phase_budget)
      if (int rc = enqueue_budget(ctx, ctx->phase_budget_dev)) return rc;
    if (overlap) {
      MARL_HIP(ctx, hipEventRecord(ctx->xev[1], ctx->stream));
      MARL_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->xev[1], 0));
    }
    // a pass on the AoS view takes the last substep's output directly (not
    // with the phase budget, which reads the quads between the phases)
    float* view = nullptr;
    if (!ctx->phase_budget && pass_uses_view(ctx, p)) {
      if (int rc = marl_ensure_staging(ctx)) return rc;
      view = ctx->staging;
    }
    if (int rc = marl_launch_field_phase(ctx, sim, true, view)) return rc;
    if (ctx->phase_budget)
      if (int rc = enqueue_budget(ctx, ctx->phase_budget_dev + kBudget)) return rc;
    MARL_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    // a small population (plain kinetics on the resident quads, no harvest or
    // phase budget) runs its cell phase in one workgroup, as the captured
    // tick does: one launch and one readback instead of some thirty launches
    // and several host round trips; ms_kinetics then covers the whole cell
    // phase and ms_schedule is 0
    const bool small = !ctx->phase_budget && ctx->kin_mode == MARL_KINETICS_PLAIN &&
                       !(ctx->surr && ctx->surr->harvesting) && ctx->n_cells <= kSmallCells * 3 / 4 &&
                       marl_choose_cell_layout(ctx);
    if (small) {
      if (!ctx->tdev) MARL_HIP(ctx, hipMalloc(&ctx->tdev, sizeof(TickDev)));
      TickDev* ts = static_cast<TickDev*>(ctx->tdev);
      MARL_HIP(ctx, hipEventRecord(ev[5], ctx->stream));
      MARL_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
      if (int rc = marl_graph_stage(ctx, ctx->stream, tick0 + t)) return rc;
      hipLaunchKernelGGL(k_g_begin, dim3(1), dim3(64), 0, ctx->stream, ts, ctx->cap);
      MARL_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
      if (int rc = marl_enqueue_cell_phase_small(ctx, sim)) return rc;
      MARL_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
      if (int rc = marl_graph_fetch(ctx, ctx->stream)) return rc;
      MARL_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
      MARL_HIP(ctx, hipEventSynchronize(ev[4]));
      int32_t finished = 0;  // a tick above kSmallCells would be finished here on the multi-launch path
      if (int rc = marl_graph_post(ctx, sim, tick0 + t, &st, &finished)) return rc;
      (void)hipEventElapsedTime(&st.ms_field, ev[0], ev[1]);
      (void)hipEventElapsedTime(&st.ms_kinetics, ev[2], ev[3]);
      (void)hipEventElapsedTime(&st.ms_fate, ev[3], ev[4]);
      (void)hipEventElapsedTime(&st.ms_tick, ev[0], ev[4]);
      continue;
    }
    {
      OnStream on(ctx, overlap ? ctx->aux : ctx->stream);
      MARL_HIP(ctx, hipEventRecord(ev[5], ctx->stream));
      if (int rc = run_schedule(ctx, p, off)) return rc;
      MARL_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
    }
    if (overlap) MARL_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev[6], 0));
    MARL_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
    if (int rc = run_levels(ctx, p, off, overlap ? ev[7] : nullptr, overlap ? ev[8] : nullptr)) return rc;
    if (ctx->phase_budget)
      if (int rc = enqueue_budget(ctx, ctx->phase_budget_dev + 2 * kBudget)) return rc;
    MARL_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
    if (int rc = run_fate(ctx, p, &st)) return rc;
    if (overlap) MARL_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev[8], 0));
    if (ctx->phase_budget) {  // against the post-compaction occupancy
      if (int rc = marl_launch_occupancy_from_cells(ctx)) return rc;
      if (int rc = enqueue_budget(ctx, ctx->phase_budget_dev + 3 * kBudget)) return rc;
    }
    MARL_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
    MARL_HIP(ctx, hipEventSynchronize(ev[4]));
    st.levels = (int32_t)off.size() - 1;
    st.n_cells = ctx->n_cells;
    (void)hipEventElapsedTime(&st.ms_field, ev[0], ev[1]);
    (void)hipEventElapsedTime(&st.ms_schedule, ev[5], ev[6]);
    (void)hipEventElapsedTime(&st.ms_kinetics, ev[2], ev[3]);
    (void)hipEventElapsedTime(&st.ms_fate, ev[3], ev[4]);
    (void)hipEventElapsedTime(&st.ms_tick, ev[0], ev[4]);
    if (overlap) (void)hipEventElapsedTime(&st.ms_transpose_back, ev[7], ev[8]);
  }
  if (stats) *stats = st;
  return MARL_OK;
}

int marl_set_graph_limits(marl_ctx* ctx, int32_t sweeps, int32_t levels) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (sweeps == 0 && levels == 0) {
    ctx->graph_auto = true;
    return MARL_OK;
  }
  if (sweeps < 1 || sweeps > kGraphSweeps || levels < 1 || levels > kGraphLevels)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "graph limits out of range");
  ctx->graph_sweeps = sweeps;
  ctx->graph_levels = levels;
  ctx->graph_auto = false;
  return MARL_OK;
}

int marl_tick_graph(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick0, int32_t n_ticks,
                    marl_tick_stats* stats, int32_t* host_finished) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (n_ticks < 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "negative tick count");
  ctx->layout_view = true;
  marl_tick_stats st{};
  int32_t finished = 0;
  for (int t = 0; t < n_ticks;) {
    if (int rc = marl_graph_prepare(ctx, sim)) return rc;
    // Overlap inside the captured tick (schedule beside the field phase,
    // transpose back beside fate) pays for large populations only: each
    // branch of the graph costs a cross-queue wait that small ticks cannot
    // hide.  (Ensembles capture their columns without it.)
    const int64_t n = ctx->n_cells;
    if (ctx->graph_overlap && n < kGraphOverlapCells * 3 / 4) ctx->graph_overlap = false;
    else if (!ctx->graph_overlap && n >= kGraphOverlapCells) ctx->graph_overlap = true;
    if (ctx->graph_overlap)
      if (int rc = marl_ensure_aux(ctx)) return rc;
    const int slot = ctx->cur;
    if (ctx->graph_key[slot] != marl_graph_key(ctx, sim))
      if (int rc = capture_tick(ctx, sim)) return rc;
    // A run of replays with no look from the host in between: in small mode
    // the device defers a tick to the host only above kSmallCells cells, a
    // population at most doubles per tick, the capacity covers kSmallCells,
    // and with an even number of buffer swaps per tick every replay is this
    // graph.  The device carries the live count and the tick index from one
    // replay to the next (k_g_finish).
    int run = 1;
    if (ctx->graph_small && !ctx->graph_flip[slot]) {
      const int64_t n0 = std::max<int64_t>(ctx->n_cells, 1);
      while (run < kSmallRun && run < n_ticks - t && (n0 << run) <= kSmallCells) ++run;
    }
    TickDev* ring = nullptr;
    if (run > 1) {
      if (!ctx->tick_ring) MARL_HIP(ctx, hipHostMalloc(&ctx->tick_ring, sizeof(TickDev) * kSmallRun, hipHostMallocDefault));
      ring = static_cast<TickDev*>(ctx->tick_ring);
    }
    if (int rc = marl_graph_stage(ctx, ctx->stream, tick0 + t)) return rc;
    MARL_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < run; ++i) {
      MARL_HIP(ctx, hipGraphLaunch(ctx->gexec[slot], ctx->stream));
      if (ring) MARL_HIP(ctx, hipMemcpyAsync(ring + i, ctx->tdev, sizeof(TickDev), hipMemcpyDeviceToHost, ctx->stream));
    }
    MARL_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!ring)
      if (int rc = marl_graph_fetch(ctx, ctx->stream)) return rc;
    MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!ring) ctx->cur ^= ctx->graph_flip[slot];
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    for (int i = 0; i < run; ++i) {
      if (ring) {
        if (ring[i].err && i + 1 < run)
          return marl_fail(ctx, MARL_ERR_TRANSFER, "captured tick deferred to the host inside a run of ticks");
        memcpy(host_tickdev(ctx), ring + i, sizeof(TickDev));
      }
      if (int rc = marl_graph_post(ctx, sim, tick0 + t + i, &st, &finished)) return rc;
      st.ms_field = ms / (float)run;
    }
    t += run;
  }
  if (stats) *stats = st;
  if (host_finished) *host_finished = finished;
  return MARL_OK;
}

int marl_set_kinetics_block(marl_ctx* ctx, int32_t block) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (block != 64 && block != 128 && block != 256)
    return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "kinetics block must be 64, 128 or 256 threads");
  ctx->kin_block = block;
  return MARL_OK;
}

int marl_get_kinetics_block(marl_ctx* ctx, int32_t* block) {
  if (!ctx || !block) return MARL_ERR_INVALID_INPUT;
  *block = ctx->kin_block;
  return MARL_OK;
}

// Times the plain cell pass on the resident field, light and population with
// each block size and keeps the fastest; the field and the cells are saved
// before the trials and restored after each, so the state is unchanged.
int marl_autotune_kinetics(marl_ctx* ctx, const marl_sim_params* sim, int32_t* block_out) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (int rc = marl_check_params(ctx, sim)) return rc;
  const int64_t n = ctx->n_cells;
  if (n == 0) {
    if (block_out) *block_out = ctx->kin_block;
    return MARL_OK;
  }
  ctx->last_sim = *sim;
  if (int rc = ensure_voxel_buffers(ctx)) return rc;
  if (int rc = marl_ensure_cell_capacity(ctx, 2 * n + 1024)) return rc;
  const size_t fbytes = ctx->nfield * sizeof(float), cbytes = (size_t)n * sizeof(marl_cell_state);
  void* save_f = nullptr;
  void* save_c = nullptr;
  MARL_HIP(ctx, hipMalloc(&save_f, fbytes));
  if (hipMalloc(&save_c, cbytes) != hipSuccess) {
    (void)hipFree(save_f);
    return marl_fail(ctx, MARL_ERR_TRANSFER, "marl_autotune_kinetics: allocation failed");
  }
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{save_f, save_c};
  float* field = ctx->field[ctx->cur];
  marl_cell_state* cells = ctx->cells[ctx->cbuf];
  MARL_HIP(ctx, hipMemcpyAsync(save_f, field, fbytes, hipMemcpyDeviceToDevice, ctx->stream));
  MARL_HIP(ctx, hipMemcpyAsync(save_c, cells, cbytes, hipMemcpyDeviceToDevice, ctx->stream));
  if (int rc = marl_launch_occupancy_from_cells(ctx)) return rc;
  const CellParams p = cell_params(ctx, sim, 0);
  const int keep = ctx->kin_block, kin_mode = ctx->kin_mode;
  ctx->kin_mode = MARL_KINETICS_PLAIN;
  const int cands[3] = {64, 128, 256};
  float best_ms = 1e30f;
  int best = keep, rc = MARL_OK;
  for (int rep = 0; rep < 2 && rc == MARL_OK; ++rep)
    for (int b : cands) {
      ctx->kin_block = b;
      int32_t levels = 0;
      hipEvent_t ev = nullptr;
      rc = hipEventCreate(&ev) == hipSuccess ? MARL_OK : MARL_ERR_TRANSFER;
      if (rc == MARL_OK) rc = run_cell_pass(ctx, p, &levels, ev);  // ev marks the end of the schedule
      if (rc == MARL_OK && hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess) rc = MARL_ERR_TRANSFER;
      if (rc == MARL_OK && hipEventSynchronize(ctx->ev1) != hipSuccess) rc = MARL_ERR_TRANSFER;
      float ms = 0.0f;
      if (rc == MARL_OK) (void)hipEventElapsedTime(&ms, ev, ctx->ev1);
      if (ev) (void)hipEventDestroy(ev);
      if (rc == MARL_OK && rep > 0 && ms < best_ms) {
        best_ms = ms;
        best = b;
      }
      if (hipMemcpyAsync(field, save_f, fbytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
          hipMemcpyAsync(cells, save_c, cbytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        rc = MARL_ERR_TRANSFER;
      if (rc != MARL_OK) break;
    }
  ctx->kin_mode = kin_mode;
  ctx->kin_block = rc == MARL_OK ? best : keep;
  MARL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (rc != MARL_OK) return marl_fail(ctx, rc, "marl_autotune_kinetics: a trial pass failed");
  if (block_out) *block_out = ctx->kin_block;
  return MARL_OK;
}

int marl_hgt_transfer(marl_ctx* ctx, const int32_t* pairs, size_t n_pairs, uint64_t seed) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  if (n_pairs == 0) return MARL_OK;
  if (!pairs) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "null pairs");
  const int64_t n = ctx->n_cells;
  std::vector<uint8_t> role((size_t)n, 0);  // 1 = donor, 2 = recipient
  for (size_t k = 0; k < n_pairs; ++k) {
    const int32_t d = pairs[2 * k], r = pairs[2 * k + 1];
    if (d < 0 || d >= n || r < 0 || r >= n) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "hgt pair out of range");
    if (d == r) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "hgt donor equals recipient");
    if (role[r] != 0) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "hgt recipient repeated or also a donor");
    role[r] = 2;
  }
  for (size_t k = 0; k < n_pairs; ++k)
    if (role[pairs[2 * k]] == 2) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "hgt recipient repeated or also a donor");
  int32_t* dp = nullptr;
  MARL_HIP(ctx, hipMalloc((void**)&dp, sizeof(int32_t) * 2 * n_pairs));
  hipError_t e = hipMemcpyAsync(dp, pairs, sizeof(int32_t) * 2 * n_pairs, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_hgt, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, ctx->stream, dp, (int)n_pairs,
                       ctx->rules[ctx->cbuf], seed);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dp);
  MARL_HIP(ctx, e);
  return MARL_OK;
}

int marl_seed_positions(const marl_dims* d, uint64_t seed, uint64_t band, int64_t count, int32_t margin,
                        int32_t z_lo, int32_t z_hi, int32_t* cell_at, uint16_t* pos_out, uint64_t* lineage_out,
                        int64_t index_base, int64_t* placed) {
  if (!d || !cell_at || !placed || count < 0 || d->x - 2 * margin <= 0 || d->y - 2 * margin <= 0)
    return MARL_ERR_INVALID_INPUT;
  const int32_t zh = std::min(z_hi, d->z);
  if (zh - z_lo <= 0 || z_lo < 0) return MARL_ERR_INVALID_INPUT;
  if (count > 0 && (!pos_out || !lineage_out)) return MARL_ERR_INVALID_INPUT;
  MarlRng rng;
  rng.init(seed, (MARL_TAG_SEEDING << 32) | band, 0);
  int64_t seeded = 0;
  for (int64_t a = 0; a < count * 8; ++a) {
    if (seeded >= count) break;
    const int x = margin + (int)rng.range((uint32_t)(d->x - 2 * margin));
    const int y = margin + (int)rng.range((uint32_t)(d->y - 2 * margin));
    const int z = z_lo + (int)rng.range((uint32_t)(zh - z_lo));
    const size_t v = ((size_t)z * d->y + y) * d->x + x;
    if (cell_at[v] >= 0) continue;
    const uint64_t lin = rng.u64();
    cell_at[v] = (int32_t)(index_base + seeded);
    pos_out[3 * seeded + 0] = (uint16_t)x;
    pos_out[3 * seeded + 1] = (uint16_t)y;
    pos_out[3 * seeded + 2] = (uint16_t)z;
    lineage_out[seeded] = lin;
    ++seeded;
  }
  *placed = seeded;
  return MARL_OK;
}

int marl_mass_budget(marl_ctx* ctx, double* out, size_t n_out) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  if (n_out < (size_t)kBudget) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "mass budget needs 40 doubles");
  if (int rc = enqueue_budget(ctx, nullptr)) return rc;
  double tmp[kBudget];
  if (int rc = d2h_sync(ctx, tmp, ctx->budget + kBudgetBlocks * kBudget, sizeof(tmp))) return rc;
  memcpy(out, tmp, sizeof(tmp));
  return MARL_OK;
}

int marl_set_phase_budget(marl_ctx* ctx, int32_t on) {
  if (!ctx) return MARL_ERR_INVALID_INPUT;
  ctx->phase_budget = on != 0;
  return MARL_OK;
}

int marl_tick_phase_budget(marl_ctx* ctx, double* out, size_t n_out) {
  if (!ctx || !out) return MARL_ERR_INVALID_INPUT;
  if (n_out < 4 * (size_t)kBudget) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "phase budget needs 160 doubles");
  if (!ctx->phase_budget_dev) return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "no tick ran with the phase budget on");
  double tmp[4 * kBudget];
  if (int rc = d2h_sync(ctx, tmp, ctx->phase_budget_dev, sizeof(tmp))) return rc;
  memcpy(out, tmp, sizeof(tmp));
  return MARL_OK;
}

}  // extern "C"
