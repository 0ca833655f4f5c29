/*
 * marl_oracle.c — CPU restatement of the reference MARL per-tick path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into the product library.
 *
 * Reference: alvertremantel/MARL-OCHRE (Rust).  Each function names the
 * file:line it restates.  Build flags (oracle/Makefile): -O3
 * -ffp-contract=off, no fast-math, so every f32 expression rounds exactly as
 * the reference's left-to-right Rust expression does (Rust applies no FMA
 * contraction).  The reference's third-party arithmetic is replaced as
 * follows, and those replacements are the parity definition of this project:
 *   - rand 0.9.2 ThreadRng (OS-seeded, crates/marl-sim/src/lib.rs:35) →
 *     Philox4x64-10 keyed by (seed, stream) with counter (draw/8, entity);
 *   - rand_distr 0.5.1 Normal → Marsaglia polar method over that stream;
 *   - libm expf (crates/marl-field/src/light.rs:45,50) → orc_expf below
 *     (Cody-Waite reduction + degree-7 Taylor polynomial).
 */
#include "marl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes.  ORC_PLANT is 0 in the oracle that tests check against.
 * tests/test_oracle_pins.py builds this file again with each nonzero value to
 * show that its independent restatements notice that mistake; no product or
 * bench path ever builds a planted variant. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT_SKIP_PARENT_HALVING 1   /* tick: the parent keeps its internals  */
#define PLANT_COUNT_SURVIVING_BIRTHS 2 /* tick: dropped births not counted    */
#define PLANT_KEEP_LAST_DUPLICATE 3   /* tick: the later duplicate birth wins  */
#define PLANT_ASCENDING_SWAP_REMOVE 4 /* tick: deaths removed low index first  */
#define PLANT_DAUGHTER_KEEPS_AGE 5    /* tick: the daughter inherits age       */
#define PLANT_STRUCTURAL_AT_RATE 6    /* mutate: rate, not rate * mult         */
#define PLANT_META_UNSCALED 7         /* mutate: meta step not * meta rate     */
#define PLANT_DONOR_AFTER_LOOP 8      /* mutate: active list read live         */
#define PLANT_SEED_LINEAGE_SWAP 9     /* seeding: lineage halves swapped       */
#define PLANT_SEED_FEWER_ATTEMPTS 10  /* seeding: count * 4 attempts           */

#define S_EXT 12
#define M_INT 16
#define R_MAX 16
#define F32_EPS 1.1920929e-07f /* f32::EPSILON = 2^-23 (cell.rs:265-266) */

/* ------------------------------------------------------------------------ */
/* Philox4x64-10 (Salmon et al., SC'11; the Random123 / numpy.random.Philox
 * bijection).  Pinned against numpy.random.Philox.random_raw in tests.     */

static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  __uint128_t p = (__uint128_t)a * (__uint128_t)b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}

void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(0xD2E7470EE14C6C93ull, c0, &hi0, &lo0);
    mulhilo64(0xCA5A826395121157ull, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw `draw` (u32) of an entity's stream: block draw/8 of the counter
 * (draw>>3, entity, 0, 0) under key (seed, stream); word (draw>>1)&3, low half
 * for even draws, high half for odd. */
uint32_t orc_draw_u32(uint64_t seed, uint64_t stream, uint64_t entity, uint64_t draw) {
  uint64_t ctr[4] = {draw >> 3, entity, 0, 0};
  uint64_t key[2] = {seed, stream};
  uint64_t out[4];
  orc_philox4x64(ctr, key, out);
  uint64_t w = out[(draw >> 1) & 3];
  return (draw & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
}

static inline float u01_from(uint32_t u) { return (float)(u >> 8) * 0x1p-24f; }
static inline uint32_t range_from(uint32_t u, uint32_t k) { return (uint32_t)(((uint64_t)u * k) >> 32); }

typedef struct { uint64_t seed, stream, entity, *draws; } rngs;
static inline uint32_t rs_u32(rngs* r) { return orc_draw_u32(r->seed, r->stream, r->entity, (*r->draws)++); }
static inline float rs_u01(rngs* r) { return u01_from(rs_u32(r)); }
static inline uint32_t rs_range(rngs* r, uint32_t k) { return range_from(rs_u32(r), k); }

/* ------------------------------------------------------------------------ */
/* expf / logf restated from the mathematics (no libm), identical in the HIP
 * library.  Pinned against libm within 2 ulp in tests.                     */

static inline float f_from_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t bits_from_f(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

float orc_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283f) return INFINITY;
  if (x < -87.33654f) return 0.0f; /* results below FLT_MIN flush to zero */
  float t = x * 1.44269504f;
  float n = floorf(t + 0.5f);
  float r = x - n * 0.693145751953125f;
  r = r - n * 1.428606765330187e-06f;
  float p = 1.98412698e-04f;   /* 1/5040 */
  p = p * r + 1.38888889e-03f; /* 1/720 */
  p = p * r + 8.33333333e-03f; /* 1/120 */
  p = p * r + 4.16666667e-02f; /* 1/24 */
  p = p * r + 1.66666667e-01f; /* 1/6 */
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int ni = (int)n;
  if (ni > 127) { p = p * 2.0f; ni -= 1; }
  return p * f_from_bits((uint32_t)(ni + 127) << 23);
}

float orc_logf(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint32_t b = bits_from_f(x);
  int e = (int)((b >> 23) & 0xff) - 127;
  float m = f_from_bits((b & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float f = (m - 1.0f) / (m + 1.0f);
  float f2 = f * f;
  float s = 1.11111111e-01f;          /* 1/9 */
  s = s * f2 + 1.42857143e-01f;       /* 1/7 */
  s = s * f2 + 2.00000000e-01f;       /* 1/5 */
  s = s * f2 + 3.33333333e-01f;       /* 1/3 */
  s = s * f2;
  float twof = 2.0f * f;
  float lm = twof + twof * s;
  float fe = (float)e;
  return fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
}

/* Standard normal by the Marsaglia polar method, scaled: mean 0 + sd*z
 * (≙ Normal::new(0.0, stddev).sample, cell.rs:408-415). */
static float rs_normal(rngs* r, float sd) {
  float u, v, s;
  do {
    u = 2.0f * rs_u01(r) - 1.0f;
    v = 2.0f * rs_u01(r) - 1.0f;
    s = u * u + v * v;
  } while (!(s < 1.0f && s > 0.0f));
  float z = u * sqrtf((-2.0f * orc_logf(s)) / s);
  return 0.0f + sd * z;
}

float orc_normal(uint64_t seed, uint64_t stream, uint64_t entity, uint64_t* draws, float sd) {
  rngs r = {seed, stream, entity, draws};
  return rs_normal(&r, sd);
}

/* Rust f32::clamp (NaN passes through). */
static inline float clampf_rs(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

/* Rust `f as u16`: truncate toward zero, saturate, NaN → 0. */
static inline uint16_t f32_as_u16(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 65535.0f) return 65535;
  return (uint16_t)f;
}

/* ------------------------------------------------------------------------ */
/* Field                                                                     */

static inline size_t vidx(const marl_dims* d, int x, int y, int z) {
  return ((size_t)z * d->y + y) * d->x + x;
}

/* Field::apply_boundary_sources — crates/marl-field/src/field.rs:231-250. */
void orc_apply_boundary_sources(float* f, const marl_dims* d, const marl_sim_params* sim) {
  for (int y = 0; y < d->y; ++y)
    for (int x = 0; x < d->x; ++x) {
      size_t b = vidx(d, x, y, 0) * S_EXT;
      f[b + 1] = fminf(f[b + 1] + sim->source_rate_oxidant, sim->c_max);
      f[b + 3] = fminf(f[b + 3] + sim->source_rate_carbon, sim->c_max);
    }
  for (int y = 0; y < d->y; ++y)
    for (int x = 0; x < d->x; ++x) {
      size_t b = vidx(d, x, y, d->z - 1) * S_EXT;
      f[b + 2] = fminf(f[b + 2] + sim->source_rate_reductant, sim->c_max);
    }
}

/* Field::diffusion_step_inner — crates/marl-field/src/field.rs:107-203
 * (second statement: crates/marl-gpu/src/shaders/field_diffuse.wgsl:55-95).
 * Parallel over z-layers like the reference's rayon par_chunks_mut (:117-120);
 * each layer writes a disjoint slice, so results do not depend on threads. */
void orc_diffusion_substep(const float* src, float* dst, const uint8_t* occ, const marl_dims* d,
                           const marl_sim_params* sim, float dt_sub, int nthreads) {
  const int X = d->x, Y = d->y, Z = d->z;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int z = 0; z < Z; ++z) {
    for (int y = 0; y < Y; ++y) {
      for (int x = 0; x < X; ++x) {
        const size_t vi = vidx(d, x, y, z);
        const size_t b = vi * S_EXT;
        if (occ && occ[vi]) {
          for (int s = 0; s < S_EXT; ++s) dst[b + s] = src[b + s];
          continue;
        }
        const float structural = src[b + 7];
        const float niche = 1.0f - sim->alpha_eps * structural / (sim->k_eps + structural);
        const int oxm = (x == 0) || (occ && occ[vi - 1]);
        const int oxp = (x >= X - 1) || (occ && occ[vi + 1]);
        const int oym = (y == 0) || (occ && occ[vi - X]);
        const int oyp = (y >= Y - 1) || (occ && occ[vi + X]);
        const int ozm = (z == 0) || (occ && occ[vi - (size_t)X * Y]);
        const int ozp = (z >= Z - 1) || (occ && occ[vi + (size_t)X * Y]);
        for (int s = 0; s < S_EXT; ++s) {
          const float c = src[b + s];
          const float dd = sim->d_voxel[s] * niche;
          const float xm = oxm ? c : src[b - S_EXT + s];
          const float xp = oxp ? c : src[b + S_EXT + s];
          const float ym = oym ? c : src[b - (size_t)X * S_EXT + s];
          const float yp = oyp ? c : src[b + (size_t)X * S_EXT + s];
          const float zm = ozm ? c : src[b - (size_t)X * Y * S_EXT + s];
          const float zp = ozp ? c : src[b + (size_t)X * Y * S_EXT + s];
          const float lap = xm + xp + ym + yp + zm + zp - 6.0f * c;
          const float decay = sim->lambda_decay[s] * c;
          const float nc = c + dt_sub * (dd * lap - decay);
          dst[b + s] = fmaxf(nc, 0.0f);
        }
      }
    }
  }
}

/* Field::diffuse_tick_with_cells — field.rs:212-217 (dt_sub at :213).  The
 * result ends in `field` (the reference swaps buffers each substep, :202). */
void orc_diffuse_tick(float* field, float* scratch, const uint8_t* occ, const marl_dims* d,
                      const marl_sim_params* sim, int nthreads) {
  const float dt_sub = sim->dt / (float)sim->diffusion_substeps;
  const size_t n = (size_t)d->x * d->y * d->z * S_EXT;
  float* a = field;
  float* b = scratch;
  for (int k = 0; k < sim->diffusion_substeps; ++k) {
    orc_diffusion_substep(a, b, occ, d, sim, dt_sub, nthreads);
    float* t = a; a = b; b = t;
  }
  if (a != field) memcpy(field, a, n * sizeof(float));
}

/* FP64 statement of the same substep (the error reference for FP32). */
void orc_diffusion_substep_f64(const double* src, double* dst, const uint8_t* occ, const marl_dims* d,
                               const marl_sim_params* sim, double dt_sub) {
  const int X = d->x, Y = d->y, Z = d->z;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int z = 0; z < Z; ++z)
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x) {
        const size_t vi = vidx(d, x, y, z);
        const size_t b = vi * S_EXT;
        if (occ && occ[vi]) {
          for (int s = 0; s < S_EXT; ++s) dst[b + s] = src[b + s];
          continue;
        }
        const double st = src[b + 7];
        const double niche = 1.0 - (double)sim->alpha_eps * st / ((double)sim->k_eps + st);
        const int oxm = (x == 0) || (occ && occ[vi - 1]);
        const int oxp = (x >= X - 1) || (occ && occ[vi + 1]);
        const int oym = (y == 0) || (occ && occ[vi - X]);
        const int oyp = (y >= Y - 1) || (occ && occ[vi + X]);
        const int ozm = (z == 0) || (occ && occ[vi - (size_t)X * Y]);
        const int ozp = (z >= Z - 1) || (occ && occ[vi + (size_t)X * Y]);
        for (int s = 0; s < S_EXT; ++s) {
          const double c = src[b + s];
          const double dd = (double)sim->d_voxel[s] * niche;
          const double xm = oxm ? c : src[b - S_EXT + s];
          const double xp = oxp ? c : src[b + S_EXT + s];
          const double ym = oym ? c : src[b - (size_t)X * S_EXT + s];
          const double yp = oyp ? c : src[b + (size_t)X * S_EXT + s];
          const double zm = ozm ? c : src[b - (size_t)X * Y * S_EXT + s];
          const double zp = ozp ? c : src[b + (size_t)X * Y * S_EXT + s];
          const double lap = xm + xp + ym + yp + zm + zp - 6.0 * c;
          const double nc = c + dt_sub * (dd * lap - (double)sim->lambda_decay[s] * c);
          dst[b + s] = nc > 0.0 ? nc : 0.0;
        }
      }
}

/* Field-only ticks in FP64: boundary sources (field.rs:231-250) then
 * sim->diffusion_substeps substeps (field.rs:107-217), n_ticks times, with a
 * fixed occupancy — the FP64 error reference for the FP32 field path. */
void orc_field_ticks_f64(double* field, double* scratch, const uint8_t* occ, const marl_dims* d,
                         const marl_sim_params* sim, int n_ticks) {
  const size_t n = (size_t)d->x * d->y * d->z * S_EXT;
  const int subs = sim->diffusion_substeps;
  const double dt_sub = subs > 0 ? (double)sim->dt / (double)subs : 0.0;
  for (int t = 0; t < n_ticks; ++t) {
    for (int y = 0; y < d->y; ++y)
      for (int x = 0; x < d->x; ++x) {
        double* f0 = field + vidx(d, x, y, 0) * S_EXT;
        double* f1 = field + vidx(d, x, y, d->z - 1) * S_EXT;
        f0[1] = fmin(f0[1] + (double)sim->source_rate_oxidant, (double)sim->c_max);
        f0[3] = fmin(f0[3] + (double)sim->source_rate_carbon, (double)sim->c_max);
        f1[2] = fmin(f1[2] + (double)sim->source_rate_reductant, (double)sim->c_max);
      }
    double* a = field;
    double* b = scratch;
    for (int k = 0; k < subs; ++k) {
      orc_diffusion_substep_f64(a, b, occ, d, sim, dt_sub);
      double* t2 = a; a = b; b = t2;
    }
    if (a != field) memcpy(field, a, n * sizeof(double));
  }
}

/* LightField::update — crates/marl-field/src/light.rs:29-59.  The constant
 * exp(-cell_absorption) is evaluated once (it does not depend on the voxel). */
void orc_light_update(float* light, const float* f, const uint8_t* occ, const marl_dims* d,
                      const marl_sim_params* sim) {
  const float e_cell = orc_expf(-sim->cell_absorption);
  const float nchem = -sim->chemical_absorption;
  for (int y = 0; y < d->y; ++y)
    for (int x = 0; x < d->x; ++x) {
      float I = sim->surface_intensity;
      for (int z = 0; z < d->z; ++z) {
        const size_t vi = vidx(d, x, y, z);
        light[vi] = I;
        if (occ && occ[vi]) I *= e_cell;
        const float absorber = f[vi * S_EXT + 4];
        I *= orc_expf(nchem * absorber);
        if (I < sim->light_floor) I = 0.0f;
      }
    }
}

void orc_light_update_f64(double* light, const double* f, const uint8_t* occ, const marl_dims* d,
                          const marl_sim_params* sim) {
  const double e_cell = exp(-(double)sim->cell_absorption);
  for (int y = 0; y < d->y; ++y)
    for (int x = 0; x < d->x; ++x) {
      double I = sim->surface_intensity;
      for (int z = 0; z < d->z; ++z) {
        const size_t vi = vidx(d, x, y, z);
        light[vi] = I;
        if (occ && occ[vi]) I *= e_cell;
        I *= exp(-(double)sim->chemical_absorption * f[vi * S_EXT + 4]);
        if (I < (double)sim->light_floor) I = 0.0;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* Cells                                                                     */

/* FACE_OFFSETS — crates/marl-sim/src/spatial.rs:8-15: +x, -x, +y, -y, +z, -z. */
static const int FACE[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

/* empty_neighbors — spatial.rs:18-43.  `cell_at` (voxel → cell index, -1 when
 * free) stands for the HashMap cell_map. */
int orc_empty_neighbors(const uint16_t pos[3], const int32_t* cell_at, const marl_dims* d, int32_t out[6]) {
  int n = 0;
  for (int k = 0; k < 6; ++k) {
    int nx = pos[0] + FACE[k][0], ny = pos[1] + FACE[k][1], nz = pos[2] + FACE[k][2];
    if (nx >= 0 && nx < d->x && ny >= 0 && ny < d->y && nz >= 0 && nz < d->z) {
      size_t v = vidx(d, nx, ny, nz);
      if (cell_at[v] < 0) out[n++] = (int32_t)v;
    }
  }
  return n;
}

/* read_neighbor_environment — spatial.rs:56-77. */
void orc_read_neighbor_environment(const uint16_t pos[3], const float* f, const int32_t* cell_at,
                                   const marl_dims* d, float ext[12]) {
  int32_t nb[6];
  int n = orc_empty_neighbors(pos, cell_at, d, nb);
  for (int s = 0; s < S_EXT; ++s) ext[s] = 0.0f;
  if (n == 0) return;
  for (int k = 0; k < n; ++k)
    for (int s = 0; s < S_EXT; ++s) ext[s] += f[(size_t)nb[k] * S_EXT + s];
  const float nf = (float)n;
  for (int s = 0; s < S_EXT; ++s) ext[s] /= nf;
}

/* apply_deltas_to_neighbors + Field::apply_deltas — spatial.rs:87-110, field.rs:57-62. */
void orc_apply_deltas_to_neighbors(const uint16_t pos[3], float* f, const int32_t* cell_at,
                                   const marl_dims* d, const float deltas[12]) {
  int32_t nb[6];
  int n = orc_empty_neighbors(pos, cell_at, d, nb);
  if (n == 0) return;
  const float nf = (float)n;
  float split[S_EXT];
  for (int s = 0; s < S_EXT; ++s) split[s] = deltas[s] / nf;
  for (int k = 0; k < n; ++k) {
    float* v = f + (size_t)nb[k] * S_EXT;
    for (int s = 0; s < S_EXT; ++s) v[s] = fmaxf(v[s] + split[s], 0.0f);
  }
}

/* CellState::tick — crates/marl-cell/src/cell.rs:193-369.  The receptor pass
 * (:206-220) writes only the unused `_activation` array and is omitted: it
 * changes no state and no output. */
int orc_cell_tick(marl_cell_state* c, const marl_ruleset* r, const float ext[12], float light,
                  const marl_sim_params* sim, float D[12]) {
  float* I = c->internal;
  const float dt = sim->dt;
  for (int s = 0; s < S_EXT; ++s) D[s] = 0.0f;

  /* transport, :223-238 */
  for (int i = 0; i < MARL_S_TRANSPORTERS; ++i) {
    const marl_transport* tp = &r->transport[i];
    const int e = tp->ext_species, k = tp->int_species;
    if (e >= S_EXT || k >= M_INT) continue;
    const float up = tp->uptake_rate * ext[e] / (1.0f + ext[e]);
    const float sec = tp->secrete_rate * I[k] / (1.0f + I[k]);
    I[k] = clampf_rs(I[k] + (up - sec) * dt, 0.0f, sim->c_max);
    D[e] += (sec - up) * dt;
  }

  /* light as pseudo-internal species 15, :243 */
  I[M_INT - 1] = light;

  /* reactions, :246-285 */
  for (int j = 0; j < R_MAX; ++j) {
    const marl_reaction* rx = &r->reactions[j];
    if (fabsf(rx->v_max) < sim->active_reaction_threshold) continue;
    if (rx->substrate == rx->product) continue;
    const int si = rx->substrate, pi = rx->product, ci = rx->catalyst;
    if (si >= M_INT || pi >= M_INT || ci >= M_INT) continue;
    const float s = I[si];
    const float cc = I[ci];
    const float st = s / (rx->k_m + s + F32_EPS);
    const float ct = sim->epsilon + cc / (rx->k_cat + cc + F32_EPS);
    float rate = rx->v_max * st * ct;
    if (rx->cofactor != 0xFF) {
      const int fi = rx->cofactor;
      if (fi < M_INT) {
        const float cof = I[fi];
        rate *= cof / (1.0f + cof);
        const float used = fminf(0.5f * rate * dt, I[fi]);
        I[fi] -= used;
      }
    }
    const float flux = fminf(rate * dt, I[si]);
    I[si] -= flux;
    I[pi] = fminf(I[pi] + flux, sim->c_max);
  }

  /* maintenance, :291-311 */
  float pm;
  if (c->prep_remaining > 0) {
    const float evolved = fmaxf(r->division_prep_ticks, 1.0f);
    const float rush = fmaxf(sim->base_division_prep - evolved, 0.0f);
    pm = sim->prep_maintenance_multiplier + rush * sim->rush_penalty_rate;
  } else {
    pm = 1.0f;
  }
  I[0] *= 1.0f - sim->lambda_maintenance * pm * dt;
  int active = 0;
  for (int j = 0; j < R_MAX; ++j)
    if (fabsf(r->reactions[j].v_max) > sim->active_reaction_threshold &&
        r->reactions[j].substrate != r->reactions[j].product)
      ++active;
  I[0] = fmaxf(I[0] - (float)active * sim->reaction_maintenance * dt, 0.0f);

  /* effectors, :314-329 */
  if (!c->quiescent) {
    for (int i = 0; i < MARL_S_EFFECTORS; ++i) {
      const marl_effector* ef = &r->effectors[i];
      const int k = ef->int_species, e = ef->ext_species;
      if (k >= M_INT || e >= S_EXT) continue;
      if (I[k] > ef->threshold) {
        const float amount = ef->rate * I[k] / (1.0f + I[k]) * dt;
        I[k] = fmaxf(I[k] - amount, 0.0f);
        D[e] += amount;
      }
    }
  }

  /* fate, :332-366 */
  const float energy = I[0];
  c->age += 1;
  const float eff_death = fmaxf(r->death_energy, sim->hard_death_floor);
  if (energy < eff_death) {
    c->prep_remaining = 0;
    return MARL_EVENT_DEATH;
  } else if (c->prep_remaining > 0) {
    c->prep_remaining -= 1;
    return c->prep_remaining == 0 ? MARL_EVENT_DIVISION : MARL_EVENT_NONE;
  } else if (energy > r->division_energy) {
    c->prep_remaining = f32_as_u16(fmaxf(r->division_prep_ticks, 1.0f));
    c->quiescent = 0;
    return MARL_EVENT_NONE;
  } else if (energy < r->quiescence_energy) {
    c->quiescent = 1;
    return MARL_EVENT_QUIESCENCE;
  }
  c->quiescent = 0;
  return MARL_EVENT_NONE;
}

/* find_empty_neighbor — spatial.rs:115-154. */
int orc_find_empty_neighbor(const uint16_t pos[3], const int32_t* cell_at, const marl_dims* d,
                            const marl_sim_params* sim, uint64_t seed, uint64_t stream, uint64_t entity,
                            uint64_t* draws, uint16_t out[3]) {
  rngs rg = {seed, stream, entity, draws};
  const int dist = sim->division_neighbor_distance;
  int cand[6][3];
  int n = 0;
  for (int k = 0; k < 6; ++k) {
    int nx = pos[0] + FACE[k][0] * dist, ny = pos[1] + FACE[k][1] * dist, nz = pos[2] + FACE[k][2] * dist;
    if (nx >= 0 && nx < d->x && ny >= 0 && ny < d->y && nz >= 0 && nz < d->z &&
        cell_at[vidx(d, nx, ny, nz)] < 0) {
      cand[n][0] = nx; cand[n][1] = ny; cand[n][2] = nz; ++n;
    }
  }
  if (n == 0) {
    for (int k = 0; k < 6; ++k) {
      int nx = pos[0] + FACE[k][0], ny = pos[1] + FACE[k][1], nz = pos[2] + FACE[k][2];
      if (nx >= 0 && nx < d->x && ny >= 0 && ny < d->y && nz >= 0 && nz < d->z &&
          cell_at[vidx(d, nx, ny, nz)] < 0) {
        cand[n][0] = nx; cand[n][1] = ny; cand[n][2] = nz; ++n;
      }
    }
    if (n == 0) return 0;
  }
  uint32_t i = rs_range(&rg, (uint32_t)n);
  out[0] = (uint16_t)cand[i][0]; out[1] = (uint16_t)cand[i][1]; out[2] = (uint16_t)cand[i][2];
  return 1;
}

/* Ruleset::mutate — crates/marl-cell/src/cell.rs:406-524, in its draw order. */
static inline void maybe_mutate(float* v, float rate, float sd, rngs* rg) {
  if (rs_u01(rg) < rate) {
    *v += rs_normal(rg, sd);
    *v = fmaxf(*v, 0.0f);
  }
}

void orc_mutate(marl_ruleset* r, const marl_sim_params* sim, uint64_t seed, uint64_t stream,
                uint64_t entity, uint64_t* draws) {
  rngs rg = {seed, stream, entity, draws};
  const float rate = r->mutation_rate;
  const float sd = sim->mutation_stddev;
#if ORC_PLANT == PLANT_STRUCTURAL_AT_RATE
  const float sr = rate;
#else
  const float sr = rate * sim->structural_mutation_rate_mult;
#endif

  for (int i = 0; i < MARL_S_RECEPTORS; ++i) { /* :419-426 */
    marl_receptor* rc = &r->receptors[i];
    maybe_mutate(&rc->k_half, rate, sd, &rg);
    maybe_mutate(&rc->n_hill, rate, sd, &rg);
    rc->n_hill = clampf_rs(rc->n_hill, sim->hill_exponent_clamp_low, sim->hill_exponent_clamp_high);
    maybe_mutate(&rc->gain, rate, sd, &rg);
  }
  for (int i = 0; i < MARL_S_TRANSPORTERS; ++i) { /* :429-438 */
    marl_transport* t = &r->transport[i];
    maybe_mutate(&t->uptake_rate, rate, sd, &rg);
    maybe_mutate(&t->secrete_rate, rate, sd, &rg);
    if (rs_u01(&rg) < sr) t->ext_species = (uint8_t)rs_range(&rg, S_EXT);
    if (rs_u01(&rg) < sr) t->int_species = (uint8_t)rs_range(&rg, M_INT);
  }
  int active[R_MAX]; /* :449-455, taken before the loop */
  int na = 0;
  for (int i = 0; i < R_MAX; ++i)
    if (fabsf(r->reactions[i].v_max) > sim->active_reaction_threshold) active[na++] = i;
  for (int i = 0; i < R_MAX; ++i) { /* :457-499 */
    marl_reaction* rx = &r->reactions[i];
    maybe_mutate(&rx->k_m, rate, sd, &rg);
    maybe_mutate(&rx->v_max, rate, sd, &rg);
    maybe_mutate(&rx->k_cat, rate, sd, &rg);
    for (int field = 0; field < 3; ++field) {
      if (rs_u01(&rg) < sr) {
        uint32_t idx = rs_range(&rg, (uint32_t)(na > 1 ? na : 1));
#if ORC_PLANT == PLANT_DONOR_AFTER_LOOP
        na = 0;
        for (int q = 0; q < R_MAX; ++q)
          if (fabsf(r->reactions[q].v_max) > sim->active_reaction_threshold) active[na++] = q;
#endif
        if ((int)idx < na) {
          int donor = active[idx];
          if (donor != i) {
            if (field == 0) rx->substrate = r->reactions[donor].substrate;
            else if (field == 1) rx->product = r->reactions[donor].product;
            else rx->catalyst = r->reactions[donor].catalyst;
          }
        } else {
          uint8_t v = (uint8_t)rs_range(&rg, M_INT);
          if (field == 0) rx->substrate = v;
          else if (field == 1) rx->product = v;
          else rx->catalyst = v;
        }
      }
    }
  }
  for (int i = 0; i < MARL_S_EFFECTORS; ++i) { /* :502-505 */
    maybe_mutate(&r->effectors[i].threshold, rate, sd, &rg);
    maybe_mutate(&r->effectors[i].rate, rate, sd, &rg);
  }
  maybe_mutate(&r->division_energy, rate, sd, &rg); /* :508-513 */
  maybe_mutate(&r->death_energy, rate, sd, &rg);
  maybe_mutate(&r->quiescence_energy, rate, sd, &rg);
  maybe_mutate(&r->division_prep_ticks, rate, sd, &rg);
  r->division_prep_ticks = fmaxf(r->division_prep_ticks, 1.0f);
  maybe_mutate(&r->hgt_propensity, rate, sd, &rg);
  if (rs_u01(&rg) < sim->meta_mutation_rate) { /* :518-523 */
#if ORC_PLANT == PLANT_META_UNSCALED
    r->mutation_rate += rs_normal(&rg, sd);
#else
    r->mutation_rate += rs_normal(&rg, sd) * sim->meta_mutation_rate;
#endif
    r->mutation_rate = clampf_rs(r->mutation_rate, sim->meta_mutation_clamp_low, sim->meta_mutation_clamp_high);
  }
}

/* Deaths removal by swap_remove in descending index order —
 * crates/marl-sim/src/lib.rs:257-267 (sequential emulation). */
int64_t orc_compact_indices(int64_t n, const int64_t* deaths, int64_t k, int64_t* perm) {
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  int64_t len = n;
  for (int64_t j = k - 1; j >= 0; --j) {
    int64_t d = deaths[j];
    perm[d] = perm[len - 1];
    --len;
  }
  return len;
}

/* The cell pass — crates/marl-sim/src/lib.rs:204-217: sequential in Vec order,
 * each cell's deltas land in the field before the next cell reads it. */
static void cell_pass_impl(const marl_dims* d, const marl_sim_params* sim, float* field, const float* light,
                           marl_cell_state* cells, const marl_ruleset* rules, int64_t n, const int32_t* cell_at,
                           uint8_t* events) {
  for (int64_t i = 0; i < n; ++i) {
    marl_cell_state* c = &cells[i];
    float ext[S_EXT], D[S_EXT];
    orc_read_neighbor_environment(c->pos, field, cell_at, d, ext);
    const float l = light[vidx(d, c->pos[0], c->pos[1], c->pos[2])];
    const int ev = orc_cell_tick(c, &rules[i], ext, l, sim, D);
    orc_apply_deltas_to_neighbors(c->pos, field, cell_at, d, D);
    events[i] = (uint8_t)ev;
  }
}

static int32_t* build_cell_at(const marl_dims* d, const marl_cell_state* cells, int64_t n) {
  const size_t nv = (size_t)d->x * d->y * d->z;
  int32_t* cell_at = (int32_t*)malloc(nv * sizeof(int32_t));
  for (size_t v = 0; v < nv; ++v) cell_at[v] = -1;
  for (int64_t i = 0; i < n; ++i) cell_at[vidx(d, cells[i].pos[0], cells[i].pos[1], cells[i].pos[2])] = (int32_t)i;
  return cell_at;
}

void orc_cell_pass(const marl_dims* d, const marl_sim_params* sim, float* field, const float* light,
                   marl_cell_state* cells, const marl_ruleset* rules, int64_t n, uint8_t* events_out) {
  int32_t* cell_at = build_cell_at(d, cells, n);
  cell_pass_impl(d, sim, field, light, cells, rules, n, cell_at, events_out);
  free(cell_at);
}

/* One tick — crates/marl-sim/src/lib.rs:160-277 (logging and dumps omitted). */
int orc_tick(const marl_dims* d, const marl_sim_params* sim, uint64_t tick, float* field, float* scratch,
             float* light, marl_cell_state* cells, marl_ruleset* rules, int64_t* n_io, int64_t capacity,
             uint8_t* events_out, orc_tick_stats* st, int nthreads) {
  const size_t nv = (size_t)d->x * d->y * d->z;
  int64_t n = *n_io;
  orc_tick_stats stats = {0, 0, 0, 0};

  orc_apply_boundary_sources(field, d, sim);                      /* :167 */
  int32_t* cell_at = build_cell_at(d, cells, n);                 /* cell_map */
  uint8_t* occ = (uint8_t*)calloc(nv, 1);                        /* :174-179 */
  for (size_t v = 0; v < nv; ++v) occ[v] = cell_at[v] >= 0;
  if (sim->diffusion_substeps > 0) orc_diffuse_tick(field, scratch, occ, d, sim, nthreads); /* :190 */
  orc_light_update(light, field, occ, d, sim);                    /* :199 */

  uint8_t* events = events_out ? events_out : (uint8_t*)malloc(n > 0 ? (size_t)n : 1);
  cell_pass_impl(d, sim, field, light, cells, rules, n, cell_at, events); /* :204-217 */

  /* fate events in cell order, :223-252 */
  marl_cell_state* bc = (marl_cell_state*)malloc(sizeof(marl_cell_state) * (size_t)(n > 0 ? n : 1));
  marl_ruleset* br = (marl_ruleset*)malloc(sizeof(marl_ruleset) * (size_t)(n > 0 ? n : 1));
  int64_t* deaths = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t nb = 0, nd = 0;
  const uint64_t stream = (ORC_TAG_DIVISION << 32) | (tick & 0xffffffffull);
  for (int64_t i = 0; i < n; ++i) {
    if (events[i] == MARL_EVENT_DIVISION) {
      uint64_t draws = 0;
      uint16_t tgt[3];
      if (orc_find_empty_neighbor(cells[i].pos, cell_at, d, sim, sim->rng_seed, stream, (uint64_t)i, &draws, tgt)) {
        marl_cell_state dau = cells[i];
        marl_ruleset dr = rules[i];
        dau.pos[0] = tgt[0]; dau.pos[1] = tgt[1]; dau.pos[2] = tgt[2];
#if ORC_PLANT != PLANT_DAUGHTER_KEEPS_AGE
        dau.age = 0;
#endif
        dau.prep_remaining = 0;
        uint32_t lo = orc_draw_u32(sim->rng_seed, stream, (uint64_t)i, draws++);
        uint32_t hi = orc_draw_u32(sim->rng_seed, stream, (uint64_t)i, draws++);
        dau.lineage_id = (uint64_t)lo | ((uint64_t)hi << 32);
        for (int k = 0; k < M_INT; ++k) {
          dau.internal[k] *= 0.5f;
#if ORC_PLANT != PLANT_SKIP_PARENT_HALVING
          cells[i].internal[k] *= 0.5f;
#endif
        }
        orc_mutate(&dr, sim, sim->rng_seed, stream, (uint64_t)i, &draws);
        bc[nb] = dau;
        br[nb] = dr;
        ++nb;
#if ORC_PLANT != PLANT_COUNT_SURVIVING_BIRTHS
        stats.divisions++;
#endif
      }
    } else if (events[i] == MARL_EVENT_DEATH) {
      deaths[nd++] = i;
      stats.deaths++;
    }
  }

  /* deaths: sort (already ascending), dedup (distinct), swap_remove descending, :257-267 */
#if ORC_PLANT == PLANT_ASCENDING_SWAP_REMOVE
  for (int64_t j0 = 0; j0 < nd; ++j0) {
    const int64_t j = j0;
#else
  for (int64_t j = nd - 1; j >= 0; --j) {
#endif
    const int64_t i = deaths[j];
    cell_at[vidx(d, cells[i].pos[0], cells[i].pos[1], cells[i].pos[2])] = -1;
    cells[i] = cells[n - 1];
    rules[i] = rules[n - 1];
    --n;
    if (i < n) cell_at[vidx(d, cells[i].pos[0], cells[i].pos[1], cells[i].pos[2])] = (int32_t)i;
  }
  /* births appended in event order, skipping taken positions, :270-277 */
  int rc = 0;
  for (int64_t j = 0; j < nb; ++j) {
    const size_t v = vidx(d, bc[j].pos[0], bc[j].pos[1], bc[j].pos[2]);
#if ORC_PLANT == PLANT_KEEP_LAST_DUPLICATE
    int later = 0; /* a later birth at the same voxel takes it instead */
    for (int64_t q = j + 1; q < nb; ++q)
      later |= bc[q].pos[0] == bc[j].pos[0] && bc[q].pos[1] == bc[j].pos[1] && bc[q].pos[2] == bc[j].pos[2];
    if (later) continue;
#endif
    if (cell_at[v] < 0) {
      if (n >= capacity) { rc = -1; break; }
      cell_at[v] = (int32_t)n;
      cells[n] = bc[j];
      rules[n] = br[j];
      ++n;
      stats.births_added++;
#if ORC_PLANT == PLANT_COUNT_SURVIVING_BIRTHS
      stats.divisions++;
#endif
    }
  }
  stats.n_cells = n;
  *n_io = n;
  if (st) *st = stats;
  free(bc); free(br); free(deaths); free(occ); free(cell_at);
  if (!events_out) free(events);
  return rc;
}

/* seed_cells — crates/marl-sim/src/seeding.rs:28-58 (positions and lineage
 * draws; the factory's fields are filled by the caller). */
int64_t orc_seed_positions(const marl_dims* d, uint64_t seed, uint64_t band, int64_t count, int32_t margin,
                           int32_t z_lo, int32_t z_hi, int32_t* cell_at, uint16_t* pos_out,
                           uint64_t* lineage_out, int64_t index_base) {
  uint64_t draws = 0;
  rngs rg = {seed, (ORC_TAG_SEEDING << 32) | band, 0, &draws};
  int64_t seeded = 0;
  const int32_t zh = z_hi < d->z ? z_hi : d->z;
#if ORC_PLANT == PLANT_SEED_FEWER_ATTEMPTS
  for (int64_t a = 0; a < count * 4; ++a) {
#else
  for (int64_t a = 0; a < count * 8; ++a) {
#endif
    if (seeded >= count) break;
    const int x = margin + (int)rs_range(&rg, (uint32_t)(d->x - 2 * margin));
    const int y = margin + (int)rs_range(&rg, (uint32_t)(d->y - 2 * margin));
    const int z = z_lo + (int)rs_range(&rg, (uint32_t)(zh - z_lo));
    const size_t v = vidx(d, x, y, z);
    if (cell_at[v] >= 0) continue;
    uint32_t lo = rs_u32(&rg), hi = rs_u32(&rg);
#if ORC_PLANT == PLANT_SEED_LINEAGE_SWAP
    { const uint32_t t = lo; lo = hi; hi = t; }
#endif
    cell_at[v] = (int32_t)(index_base + seeded);
    pos_out[3 * seeded + 0] = (uint16_t)x;
    pos_out[3 * seeded + 1] = (uint16_t)y;
    pos_out[3 * seeded + 2] = (uint16_t)z;
    lineage_out[seeded] = (uint64_t)lo | ((uint64_t)hi << 32);
    ++seeded;
  }
  return seeded;
}

/* ------------------------------------------------------------------------ */
/* hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27): a donor's active
 * reaction (|v_max| > 1e-9), drawn uniformly, overwrites the recipient's first
 * inactive slot (|v_max| < 1e-9), or a drawn slot when none is inactive.      */
void orc_hgt_transfer(marl_ruleset* rules, const int32_t* pairs, int64_t n_pairs, uint64_t seed) {
  for (int64_t k = 0; k < n_pairs; ++k) {
    const marl_ruleset* d = &rules[pairs[2 * k]];
    marl_ruleset* r = &rules[pairs[2 * k + 1]];
    int active[MARL_R_MAX], na = 0;
    for (int q = 0; q < MARL_R_MAX; ++q)
      if (fabsf(d->reactions[q].v_max) > 1e-9f) active[na++] = q;
    if (na == 0) continue;
    uint64_t draws = 0;
    rngs rg = {seed, ORC_TAG_HGT << 32, (uint64_t)k, &draws};
    const int src = active[rs_range(&rg, (uint32_t)na)];
    int tgt = -1;
    for (int q = 0; q < MARL_R_MAX && tgt < 0; ++q)
      if (fabsf(r->reactions[q].v_max) < 1e-9f) tgt = q;
    if (tgt < 0) tgt = (int)rs_range(&rg, MARL_R_MAX);
    r->reactions[tgt] = d->reactions[src];
  }
}
