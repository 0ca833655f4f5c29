// kinetics.h — per-cell device functions shared by the plain kinetics kernel,
// the pair harvester and the surrogate's fallback path.
//
// Reference: crates/marl-sim/src/spatial.rs:56-110 (environment, deltas),
// crates/marl-cell/src/cell.rs:193-369 (CellState::tick).
#pragma once

#include "marl_internal.h"

struct CellParams {
  float dt, epsilon, c_max, lambda_maintenance, hard_death_floor, reaction_maintenance;
  float base_division_prep, prep_maintenance_multiplier, rush_penalty_rate, active_reaction_threshold;
  float mutation_stddev, structural_mutation_rate_mult, meta_mutation_rate, meta_lo, meta_hi, hill_lo, hill_hi;
  int32_t div_dist;
  int32_t X, Y, Z;
  uint64_t seed, stream;
  uint64_t fplane;  // floats between species-quad planes of the resident field (4 x voxels)
};

__host__ __device__ inline float clamp_rs(float x, float lo, float hi) {  // Rust f32::clamp
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

__device__ inline uint16_t f32_as_u16(float f) {  // Rust `as u16`: truncate, saturate, NaN -> 0
  if (!(f > 0.0f)) return 0;
  if (f >= 65535.0f) return 65535;
  return (uint16_t)f;
}

__device__ __forceinline__ marl_cell_state load_cell(const marl_cell_state* c) {
  constexpr int N = sizeof(marl_cell_state) / 16;
  const uint4* s = reinterpret_cast<const uint4*>(c);
  uint4 buf[N];
#pragma unroll
  for (int k = 0; k < N; ++k) buf[k] = s[k];
  marl_cell_state out;
  __builtin_memcpy(&out, buf, sizeof(out));
  return out;
}

// The parameters the tick reads from a cell's ruleset, through accessors so
// that cell_kinetics does not depend on the record layout.
struct AosRules {
  const marl_ruleset* r;
  __device__ __forceinline__ marl_transport transport(int q) const { return r->transport[q]; }
  __device__ __forceinline__ marl_reaction reaction(int q) const { return r->reactions[q]; }
  __device__ __forceinline__ marl_effector effector(int q) const { return r->effectors[q]; }
  // (division, death, quiescence energies, division prep ticks)
  __device__ __forceinline__ float4 fate() const {
    return make_float4(r->division_energy, r->death_energy, r->quiescence_energy, r->division_prep_ticks);
  }
};

// A voxel's species quad k (species 4k..4k+3) in the field the cell pass works
// on: the AoS view (one 48-byte record per voxel, quads 16 bytes apart) or,
// with QUADS, the resident species quads in place (planes p.fplane floats
// apart, 16 bytes per voxel in each).
template <bool QUADS>
__device__ __forceinline__ float4* field_quad(float* field, size_t w, int k, const CellParams& p) {
  return QUADS ? reinterpret_cast<float4*>(field + (size_t)k * p.fplane + w * 4)
               : reinterpret_cast<float4*>(field + w * MARL_S_EXT) + k;
}

// Mean environment over the free face neighbours (spatial.rs:56-77; faces in
// FACE_OFFSETS order +x,-x,+y,-y,+z,-z, spatial.rs:18-43): all six face
// records are requested at once (non-free faces read the cell's own voxel and
// are ignored), the sum runs over the free faces in face order, then divides;
// zeros when enclosed.
template <bool QUADS>
__device__ inline void gather_env6(const float* __restrict__ field, size_t v, uint32_t m, const CellParams& p,
                                   float* E) {
  const size_t XY = (size_t)p.X * p.Y;
  const int64_t off[6] = {1, -1, (int64_t)p.X, -(int64_t)p.X, (int64_t)XY, -(int64_t)XY};
  float* f = const_cast<float*>(field);
  float4 a[6], b[6], cc[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const size_t w = ((m >> k) & 1u) ? (size_t)((int64_t)v + off[k]) : v;
    a[k] = *field_quad<QUADS>(f, w, 0, p);
    b[k] = *field_quad<QUADS>(f, w, 1, p);
    cc[k] = *field_quad<QUADS>(f, w, 2, p);
  }
  float acc[MARL_S_EXT];
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) acc[s] = 0.0f;
  int n = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if ((m >> k) & 1u) {
      ++n;
      acc[0] += a[k].x; acc[1] += a[k].y; acc[2] += a[k].z; acc[3] += a[k].w;
      acc[4] += b[k].x; acc[5] += b[k].y; acc[6] += b[k].z; acc[7] += b[k].w;
      acc[8] += cc[k].x; acc[9] += cc[k].y; acc[10] += cc[k].z; acc[11] += cc[k].w;
    }
  }
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) E[s] = n ? acc[s] / (float)n : 0.0f;
}

// gather_env6 that also keeps the six face records (18 quads) for the
// scatter: within a level no other cell touches these voxels, so the values
// the scatter updates are exactly the ones gathered, and the scatter then
// writes without reading the field again.
template <bool QUADS>
__device__ inline void gather_env6_keep(const float* __restrict__ field, size_t v, uint32_t m, const CellParams& p,
                                        float* E, float4 (&keep)[18]) {
  const size_t XY = (size_t)p.X * p.Y;
  const int64_t off[6] = {1, -1, (int64_t)p.X, -(int64_t)p.X, (int64_t)XY, -(int64_t)XY};
  float* f = const_cast<float*>(field);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const size_t w = ((m >> k) & 1u) ? (size_t)((int64_t)v + off[k]) : v;
    keep[3 * k] = *field_quad<QUADS>(f, w, 0, p);
    keep[3 * k + 1] = *field_quad<QUADS>(f, w, 1, p);
    keep[3 * k + 2] = *field_quad<QUADS>(f, w, 2, p);
  }
  float acc[MARL_S_EXT];
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) acc[s] = 0.0f;
  int n = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if ((m >> k) & 1u) {
      ++n;
      const float4 a = keep[3 * k], b = keep[3 * k + 1], cc = keep[3 * k + 2];
      acc[0] += a.x; acc[1] += a.y; acc[2] += a.z; acc[3] += a.w;
      acc[4] += b.x; acc[5] += b.y; acc[6] += b.z; acc[7] += b.w;
      acc[8] += cc.x; acc[9] += cc.y; acc[10] += cc.z; acc[11] += cc.w;
    }
  }
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) E[s] = n ? acc[s] / (float)n : 0.0f;
}

// scatter_deltas6 from the records gather_env6_keep kept (same arithmetic).
template <bool QUADS>
__device__ inline void scatter_from_kept(float* __restrict__ field, size_t v, uint32_t m, const CellParams& p,
                                         const float* D, const float4 (&keep)[18]) {
  const int n = __builtin_popcount(m);
  if (n == 0) return;
  const size_t XY = (size_t)p.X * p.Y;
  const int64_t off[6] = {1, -1, (int64_t)p.X, -(int64_t)p.X, (int64_t)XY, -(int64_t)XY};
  const float nf = (float)n;
  float split[MARL_S_EXT];
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) split[s] = D[s] / nf;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!((m >> k) & 1u)) continue;
    const size_t w = (size_t)((int64_t)v + off[k]);
    float4 a = keep[3 * k], b = keep[3 * k + 1], cc = keep[3 * k + 2];
    a.x = fmaxf(a.x + split[0], 0.0f); a.y = fmaxf(a.y + split[1], 0.0f);
    a.z = fmaxf(a.z + split[2], 0.0f); a.w = fmaxf(a.w + split[3], 0.0f);
    b.x = fmaxf(b.x + split[4], 0.0f); b.y = fmaxf(b.y + split[5], 0.0f);
    b.z = fmaxf(b.z + split[6], 0.0f); b.w = fmaxf(b.w + split[7], 0.0f);
    cc.x = fmaxf(cc.x + split[8], 0.0f); cc.y = fmaxf(cc.y + split[9], 0.0f);
    cc.z = fmaxf(cc.z + split[10], 0.0f); cc.w = fmaxf(cc.w + split[11], 0.0f);
    *field_quad<QUADS>(field, w, 0, p) = a;
    *field_quad<QUADS>(field, w, 1, p) = b;
    *field_quad<QUADS>(field, w, 2, p) = cc;
  }
}

// Split deltas into the free face neighbours with the max(0) clamp
// (spatial.rs:87-110, field.rs:57-62), lost when enclosed.  Each free face is
// a different voxel, so the order of the updates does not matter.
template <bool QUADS>
__device__ inline void scatter_deltas6(float* __restrict__ field, size_t v, uint32_t m, const CellParams& p,
                                       const float* D) {
  const int n = __builtin_popcount(m);
  if (n == 0) return;
  const size_t XY = (size_t)p.X * p.Y;
  const int64_t off[6] = {1, -1, (int64_t)p.X, -(int64_t)p.X, (int64_t)XY, -(int64_t)XY};
  const float nf = (float)n;
  float split[MARL_S_EXT];
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) split[s] = D[s] / nf;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!((m >> k) & 1u)) continue;
    const size_t w = (size_t)((int64_t)v + off[k]);
    float4* qa = field_quad<QUADS>(field, w, 0, p);
    float4* qb = field_quad<QUADS>(field, w, 1, p);
    float4* qc = field_quad<QUADS>(field, w, 2, p);
    float4 a = *qa, b = *qb, cc = *qc;
    a.x = fmaxf(a.x + split[0], 0.0f); a.y = fmaxf(a.y + split[1], 0.0f);
    a.z = fmaxf(a.z + split[2], 0.0f); a.w = fmaxf(a.w + split[3], 0.0f);
    b.x = fmaxf(b.x + split[4], 0.0f); b.y = fmaxf(b.y + split[5], 0.0f);
    b.z = fmaxf(b.z + split[6], 0.0f); b.w = fmaxf(b.w + split[7], 0.0f);
    cc.x = fmaxf(cc.x + split[8], 0.0f); cc.y = fmaxf(cc.y + split[9], 0.0f);
    cc.z = fmaxf(cc.z + split[10], 0.0f); cc.w = fmaxf(cc.w + split[11], 0.0f);
    *qa = a; *qb = b; *qc = cc;
  }
}

// Phase 5 (cell.rs:332-366) from the post-effector energy.
// fate = (division, death, quiescence energies, division prep ticks).
__device__ inline uint8_t cell_fate(marl_cell_state& c, const float4 fate, float energy, const CellParams& p) {
  c.age += 1;
  const float eff_death = fmaxf(fate.y, p.hard_death_floor);
  if (energy < eff_death) {
    c.prep_remaining = 0;
    return MARL_EVENT_DEATH;
  } else if (c.prep_remaining > 0) {
    c.prep_remaining -= 1;
    return c.prep_remaining == 0 ? MARL_EVENT_DIVISION : MARL_EVENT_NONE;
  } else if (energy > fate.x) {
    c.prep_remaining = f32_as_u16(fmaxf(fate.w, 1.0f));
    c.quiescent = 0;
    return MARL_EVENT_NONE;
  } else if (energy < fate.z) {
    c.quiescent = 1;
    return MARL_EVENT_QUIESCENCE;
  }
  c.quiescent = 0;
  return MARL_EVENT_NONE;
}

// Per-cell arrays indexed by data (species numbers taken from the ruleset).
// PtrArr views a thread's LDS slot.  RegArr<N> keeps the array in registers:
// a read is a select tree over the index bits and a write one select per
// element, so a data-dependent index costs a few dependent VALU operations
// instead of an LDS round trip (the cell update is a chain of such accesses).
struct PtrArr {
  float* p;
  __device__ __forceinline__ float get(int i) const { return p[i]; }
  __device__ __forceinline__ void set(int i, float x) { p[i] = x; }
};

template <int N>
struct RegArr {
  static_assert(N <= 16, "RegArr size");
  float v[N];
  __device__ __forceinline__ float get(int i) const {
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = k < N ? v[k] : 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = (i & 1) ? a[2 * k + 1] : a[2 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = (i & 2) ? a[2 * k + 1] : a[2 * k];
#pragma unroll
    for (int k = 0; k < 2; ++k) a[k] = (i & 4) ? a[2 * k + 1] : a[2 * k];
    return (i & 8) ? a[1] : a[0];
  }
  __device__ __forceinline__ void set(int i, float x) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = (i == k) ? x : v[k];
  }
};

// The ruleset's transporter, reaction and effector records are read in
// batches of kRuleBatch: a batch's loads are issued together, before any of its
// records is used, so a cell's update waits for one memory round trip per
// batch instead of one per record (the rules are read-only during a pass).
#ifndef MARL_KIN_BATCH
#define MARL_KIN_BATCH 16
#endif
constexpr int kRuleBatch = MARL_KIN_BATCH;

template <int N, typename T, typename Load, typename Body>
__device__ __forceinline__ void batched_loop(Load load, Body body) {
  constexpr int B = kRuleBatch < 1 ? 1 : (kRuleBatch < N ? kRuleBatch : N);
  static_assert(N % B == 0, "batch divides the record count");
#pragma unroll
  for (int q0 = 0; q0 < N; q0 += B) {
    T buf[B];
#pragma unroll
    for (int k = 0; k < B; ++k) buf[k] = load(q0 + k);
#pragma unroll
    for (int k = 0; k < B; ++k) body(q0 + k, buf[k]);
  }
}

// Phases 2-4 of CellState::tick (cell.rs:223-329) on I[16] (updated in place)
// with environment E[12], accumulating exchange fluxes into D[12].  The
// receptor pass (:206-220) only writes an unused array and is omitted.
template <typename R, typename AI, typename AE, typename AD>
__device__ inline void cell_kinetics(const marl_cell_state& c, const R& r, AI& I, const AE& E, AD& D, float lt,
                                     const CellParams& p) {
  const float dt = p.dt;
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) D.set(s, 0.0f);
  // transport (cell.rs:223-238)
  batched_loop<MARL_S_TRANSPORTERS, marl_transport>([&](int q) { return r.transport(q); },
                                                       [&](int, const marl_transport& tp) {
    const int e = tp.ext_species, k = tp.int_species;
    if (e >= MARL_S_EXT || k >= MARL_M_INT) return;
    const float ex = E.get(e);
    const float ik = I.get(k);
    const float up = tp.uptake_rate * ex / (1.0f + ex);
    const float sec = tp.secrete_rate * ik / (1.0f + ik);
    I.set(k, clamp_rs(ik + (up - sec) * dt, 0.0f, p.c_max));
    D.set(e, D.get(e) + (sec - up) * dt);
  });
  I.set(MARL_M_INT - 1, lt);  // :243
  // reactions (cell.rs:246-285)
  int active = 0;
  batched_loop<MARL_R_MAX, marl_reaction>([&](int q) { return r.reaction(q); },
                                             [&](int, const marl_reaction& rx) {
    if (fabsf(rx.v_max) > p.active_reaction_threshold && rx.substrate != rx.product) ++active;
    if (fabsf(rx.v_max) < p.active_reaction_threshold) return;
    if (rx.substrate == rx.product) return;
    const int si = rx.substrate, pi = rx.product, ci = rx.catalyst;
    if (si >= MARL_M_INT || pi >= MARL_M_INT || ci >= MARL_M_INT) return;
    const float s = I.get(si);
    const float cc = I.get(ci);
    const float st = s / (rx.k_m + s + 1.1920929e-07f);
    const float ct = p.epsilon + cc / (rx.k_cat + cc + 1.1920929e-07f);
    float rate = rx.v_max * st * ct;
    if (rx.cofactor != 0xFF) {
      const int fi = rx.cofactor;
      if (fi < MARL_M_INT) {
        const float cof = I.get(fi);
        rate *= cof / (1.0f + cof);
        const float used = fminf(0.5f * rate * dt, I.get(fi));
        I.set(fi, I.get(fi) - used);
      }
    }
    const float flux = fminf(rate * dt, I.get(si));
    I.set(si, I.get(si) - flux);
    I.set(pi, fminf(I.get(pi) + flux, p.c_max));
  });
  // maintenance (cell.rs:291-311)
  float pm = 1.0f;
  if (c.prep_remaining > 0) {
    const float evolved = fmaxf(r.fate().w, 1.0f);
    const float rush = fmaxf(p.base_division_prep - evolved, 0.0f);
    pm = p.prep_maintenance_multiplier + rush * p.rush_penalty_rate;
  }
  float e0 = I.get(0);
  e0 *= 1.0f - p.lambda_maintenance * pm * dt;
  e0 = fmaxf(e0 - (float)active * p.reaction_maintenance * dt, 0.0f);
  I.set(0, e0);
  // effectors (cell.rs:314-329)
  if (!c.quiescent) {
    batched_loop<MARL_S_EFFECTORS, marl_effector>([&](int q) { return r.effector(q); },
                                                     [&](int, const marl_effector& ef) {
      const int k = ef.int_species, e = ef.ext_species;
      if (k >= MARL_M_INT || e >= MARL_S_EXT) return;
      const float ik = I.get(k);
      if (ik > ef.threshold) {
        const float amount = ef.rate * ik / (1.0f + ik) * dt;
        I.set(k, fmaxf(ik - amount, 0.0f));
        D.set(e, D.get(e) + amount);
      }
    });
  }
}

// ---------------------------------------------------------------------------
// Surrogate inputs and targets.
//
// The surrogate predicts the residual between the exact update (cell_kinetics:
// transporters, reactions and effectors applied one after another, each seeing
// the previous ones' results, with their clamps) and one explicit step of the
// same rate laws evaluated from the pre-tick state and routed through the
// cell's own wiring (explicit_step).  The explicit step carries every cell's
// wiring into the features, so any wiring is eligible; the network learns what
// the sequential order and the clamps change.
// Features (80): internal[0..14], quiescent, environment[12], light, in-prep,
// the explicit internal deltas [0..14], the explicit exchange fluxes [12], each
// reaction's explicit flux relative to its substrate [16], then 7 zeros.
// Targets (28): exact minus explicit internal deltas [16] (slot 15, which the
// tick overwrites with the voxel's light, is not learned) and exact minus
// explicit exchange fluxes [12], each divided by its output's scale, the
// output's gross explicit activity (the sum of the magnitudes of the explicit
// terms that touch it) plus kGrossEps, and clipped to +-kTargClip.  A slow
// output then gets a correction as small as its own activity, so the
// network's error cannot swamp outputs the explicit step already gets right.
// kGrossEps keeps the f32 rounding of the residual itself (about 6e-8 of an
// internal of order one) from turning the scaled targets of the slowest
// outputs into noise.
// An output that no explicit term touches has scale 0: its target is 0 and it
// gets no correction.  What the clip cuts off (an output reached only through
// a chain of reactions within the tick) stays an error of the prediction.
constexpr int kFeat = 80;
constexpr int kTarg = 28;
constexpr int kFeatExplicit = 30;  // first explicit feature
constexpr float kGrossEps = 1e-5f;
constexpr float kTargClip = 4.0f;

// Target scales from the gross activities: s[0..15] internals, s[16..27]
// exchange fluxes.
__device__ __forceinline__ void target_scales(const float* gI, const float* gD, float* s) {
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) s[k] = gI[k] > 0.0f ? gI[k] + kGrossEps : 0.0f;
#pragma unroll
  for (int e = 0; e < MARL_S_EXT; ++e) s[16 + e] = gD[e] > 0.0f ? gD[e] + kGrossEps : 0.0f;
}

// A residual over its scale, clipped (0 where the scale is 0).
__device__ __forceinline__ float scaled_target(float resid, float scale) {
  return scale > 0.0f ? clamp_rs(resid / scale, -kTargClip, kTargClip) : 0.0f;
}

// One explicit step of cell.rs:223-329's rate laws from the pre-tick
// internals I0 and environment E: transport, reactions (light in slot 15, as
// the reactions see it), maintenance on the pre-tick energy and effectors, all
// evaluated at the same state and summed into dI[16] / dD[12] without clamps;
// frel[q] = reaction q's flux / (its substrate + 1e-6); gI[16] / gD[12] = the
// gross activity of each output (sum of the magnitudes of its terms).
template <typename R, typename AI, typename AE, typename AQ, typename AD, typename AG, typename AH>
__device__ inline void explicit_step(const marl_cell_state& c, const R& r, const AI& I0, const AE& E, AQ& dI, AD& dD,
                                     float lt, const CellParams& p, float* frel, AG& gI, AH& gD) {
  const float dt = p.dt;
#pragma unroll
  for (int k = 0; k < MARL_M_INT; ++k) {
    dI.set(k, 0.0f);
    gI.set(k, 0.0f);
  }
#pragma unroll
  for (int e = 0; e < MARL_S_EXT; ++e) {
    dD.set(e, 0.0f);
    gD.set(e, 0.0f);
  }
#pragma unroll
  for (int q = 0; q < MARL_S_TRANSPORTERS; ++q) {
    const marl_transport tp = r.transport(q);
    const int e = tp.ext_species, k = tp.int_species;
    if (e >= MARL_S_EXT || k >= MARL_M_INT) continue;
    const float ex = E.get(e), ik = I0.get(k);
    const float up = tp.uptake_rate * ex / (1.0f + ex);
    const float sec = tp.secrete_rate * ik / (1.0f + ik);
    dI.set(k, dI.get(k) + (up - sec) * dt);
    dD.set(e, dD.get(e) + (sec - up) * dt);
    const float gross = (fabsf(up) + fabsf(sec)) * dt;
    gI.set(k, gI.get(k) + gross);
    gD.set(e, gD.get(e) + gross);
  }
  auto val = [&](int i) { return i == MARL_M_INT - 1 ? lt : I0.get(i); };
  int active = 0;
#pragma unroll
  for (int q = 0; q < MARL_R_MAX; ++q) {
    const marl_reaction rx = r.reaction(q);
    frel[q] = 0.0f;
    if (fabsf(rx.v_max) > p.active_reaction_threshold && rx.substrate != rx.product) ++active;
    if (fabsf(rx.v_max) < p.active_reaction_threshold || rx.substrate == rx.product) continue;
    const int si = rx.substrate, pi = rx.product, ci = rx.catalyst;
    if (si >= MARL_M_INT || pi >= MARL_M_INT || ci >= MARL_M_INT) continue;
    const float sv = val(si), cv = val(ci);
    float rate = rx.v_max * (sv / (rx.k_m + sv + 1.1920929e-07f)) *
                 (p.epsilon + cv / (rx.k_cat + cv + 1.1920929e-07f));
    if (rx.cofactor != 0xFF && rx.cofactor < MARL_M_INT) {
      const int fi = rx.cofactor;
      const float cof = val(fi);
      rate *= cof / (1.0f + cof);
      dI.set(fi, dI.get(fi) - 0.5f * rate * dt);
      gI.set(fi, gI.get(fi) + fabsf(0.5f * rate * dt));
    }
    const float flux = rate * dt;
    dI.set(si, dI.get(si) - flux);
    dI.set(pi, dI.get(pi) + flux);
    gI.set(si, gI.get(si) + fabsf(flux));
    gI.set(pi, gI.get(pi) + fabsf(flux));
    frel[q] = flux / (sv + 1e-6f);
  }
  float pm = 1.0f;
  if (c.prep_remaining > 0) {
    const float evolved = fmaxf(r.fate().w, 1.0f);
    pm = p.prep_maintenance_multiplier + fmaxf(p.base_division_prep - evolved, 0.0f) * p.rush_penalty_rate;
  }
  const float decay = p.lambda_maintenance * pm * dt * I0.get(0), upkeep = (float)active * p.reaction_maintenance * dt;
  dI.set(0, dI.get(0) - decay - upkeep);
  gI.set(0, gI.get(0) + fabsf(decay) + fabsf(upkeep));
  if (!c.quiescent) {
#pragma unroll
    for (int q = 0; q < MARL_S_EFFECTORS; ++q) {
      const marl_effector ef = r.effector(q);
      const int k = ef.int_species, e = ef.ext_species;
      if (k >= MARL_M_INT || e >= MARL_S_EXT) continue;
      const float ik = I0.get(k);
      if (ik > ef.threshold) {
        const float amount = ef.rate * ik / (1.0f + ik) * dt;
        dI.set(k, dI.get(k) - amount);
        dD.set(e, dD.get(e) + amount);
        gI.set(k, gI.get(k) + fabsf(amount));
        gD.set(e, gD.get(e) + fabsf(amount));
      }
    }
  }
}

__device__ inline void make_features(const float* I_pre, const marl_cell_state& c, const float* E, float lt,
                                     const float* dI, const float* dD, const float* frel, float* f) {
#pragma unroll
  for (int k = 0; k < 15; ++k) f[k] = I_pre[k];
  f[15] = c.quiescent ? 1.0f : 0.0f;
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) f[16 + s] = E[s];
  f[28] = lt;
  f[29] = c.prep_remaining > 0 ? 1.0f : 0.0f;
#pragma unroll
  for (int k = 0; k < 15; ++k) f[kFeatExplicit + k] = dI[k];
#pragma unroll
  for (int e = 0; e < MARL_S_EXT; ++e) f[kFeatExplicit + 15 + e] = dD[e];
#pragma unroll
  for (int q = 0; q < MARL_R_MAX; ++q) f[kFeatExplicit + 27 + q] = frel[q];
#pragma unroll
  for (int k = kFeatExplicit + 43; k < kFeat; ++k) f[k] = 0.0f;
}

// 64-bit FNV-1a over the ruleset's integer wiring (reaction species,
// transporter and effector species): cells with the same hash share a
// reaction topology.
template <typename R>
__device__ inline uint64_t topology_hash(const R& r) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint32_t b) {
    h ^= b;
    h *= 1099511628211ull;
  };
#pragma unroll 4
  for (int q = 0; q < MARL_R_MAX; ++q) {
    const marl_reaction rx = r.reactions[q];
    mix(rx.substrate | (rx.product << 8) | (rx.catalyst << 16) | ((uint32_t)rx.cofactor << 24));
  }
#pragma unroll
  for (int q = 0; q < MARL_S_TRANSPORTERS; ++q) mix(r.transport[q].ext_species | (r.transport[q].int_species << 8));
#pragma unroll
  for (int q = 0; q < MARL_S_EFFECTORS; ++q) mix(r.effectors[q].int_species | (r.effectors[q].ext_species << 8));
  return h;
}

// Pair harvest: sampled cells of the plain pass append (features, targets).
struct HarvestSink {
  float* feat;        // [cap][kFeat] (class slots zero; class from hash at training time)
  float* targ;        // [cap][kTarg]: internal deltas [16], exchange fluxes [12]
  uint64_t* hash;     // [cap] topology hash
  unsigned long long* count;
  int64_t cap;
  uint64_t threshold;  // sample if hash32(tick, cell) < threshold (2^32: every cell)
  uint64_t tick;
};

__device__ inline uint32_t sample_hash(uint64_t tick, uint32_t i) {
  uint64_t z = tick * 0x9E3779B97F4A7C15ull + i;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)(z ^ (z >> 31));
}
