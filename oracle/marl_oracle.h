/*
 * marl_oracle.h — CPU restatement of the reference MARL tick (test infrastructure).
 *
 * This is the oracle that the HIP path is checked against.  It is used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product
 * library never links or calls it.  Every function cites the reference
 * (alvertremantel/MARL-OCHRE) file:line it restates.  Arithmetic is f32 in the
 * reference's operator order, built with -ffp-contract=off and no fast-math, so
 * that IEEE-754 rounding is fixed; the *_f64 functions are double-precision
 * versions of the stencil and light scan used to state the FP32 error.
 *
 * Pinning (what checks each function; see tests/):
 *   orc_philox4x64        numpy.random.Philox golden vectors (tests/golden)
 *   orc_expf / orc_logf   libm expf/logf within 2 ulp over sweeps
 *   stencil / boundary    reference tests field.rs:257-312 and the six cases of
 *                         crates/marl-gpu/tests/gpu_diffusion.rs, closed forms
 *                         (uniform-field pure decay, impulse weights, mirror
 *                         symmetry, FP64 mass conservation), FP64 brute force
 *   light                 FP64 closed form, geometric decay, floor switch
 *   cell tick             hand-derived known answers, invariants, fate truth table
 *   neighbours / fate /
 *   compaction            brute-force sequential emulation (Python) on tiny grids;
 *                         the tick's fate, birth, death and compaction block
 *                         against a literal Python restatement of
 *                         crates/marl-sim/src/lib.rs:220-277 (tests/restate.py)
 *   mutate                a draw-order Python restatement of cell.rs:406-524
 *                         (structural mult != 1, meta mutation), the structural
 *                         rate rate * mult and the meta step's spread measured,
 *                         the normal sampler's moments and KS distance
 *   seeding               a Python restatement of seeding.rs:28-58
 *   (each of the last three also catches mistakes planted with ORC_PLANT);
 *   the draw streams themselves stay unpinned against the reference, whose
 *   rand 0.9.2 ThreadRng is OS-seeded (crates/marl-sim/src/lib.rs:35)
 *   hgt transfer          slot rules (first inactive slot, else a drawn one),
 *                         donor untouched, the v_max thresholds, the pick's
 *                         uniformity over active reactions (chi-square) and its
 *                         widening-multiply mapping of the drawn u32
 */
#ifndef MARL_ORACLE_H
#define MARL_ORACLE_H

#include <stdint.h>
#include <stddef.h>
#include "../include/marl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RNG stream tags (key[1] = tag << 32 | sub-stream). */
#define ORC_TAG_SEEDING 1ull
#define ORC_TAG_DIVISION 2ull
#define ORC_TAG_HGT 3ull

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
uint32_t orc_draw_u32(uint64_t seed, uint64_t stream, uint64_t entity, uint64_t draw);
float orc_expf(float x);
float orc_logf(float x);

/* field (AoS) */
void orc_apply_boundary_sources(float* field, const marl_dims* d, const marl_sim_params* sim);
void orc_diffusion_substep(const float* src, float* dst, const uint8_t* occ, const marl_dims* d,
                           const marl_sim_params* sim, float dt_sub, int nthreads);
void orc_diffuse_tick(float* field, float* scratch, const uint8_t* occ, const marl_dims* d,
                      const marl_sim_params* sim, int nthreads);
void orc_diffusion_substep_f64(const double* src, double* dst, const uint8_t* occ, const marl_dims* d,
                               const marl_sim_params* sim, double dt_sub);
/* n_ticks field-only ticks (boundary sources + substeps) in FP64, fixed occupancy. */
void orc_field_ticks_f64(double* field, double* scratch, const uint8_t* occ, const marl_dims* d,
                         const marl_sim_params* sim, int n_ticks);
void orc_light_update(float* light, const float* field, const uint8_t* occ, const marl_dims* d,
                      const marl_sim_params* sim);
void orc_light_update_f64(double* light, const double* field, const uint8_t* occ, const marl_dims* d,
                          const marl_sim_params* sim);

/* cells */
int orc_empty_neighbors(const uint16_t pos[3], const int32_t* cell_at, const marl_dims* d, int32_t out_vox[6]);
void orc_read_neighbor_environment(const uint16_t pos[3], const float* field, const int32_t* cell_at,
                                   const marl_dims* d, float ext[12]);
void orc_apply_deltas_to_neighbors(const uint16_t pos[3], float* field, const int32_t* cell_at,
                                   const marl_dims* d, const float deltas[12]);
int orc_cell_tick(marl_cell_state* c, const marl_ruleset* r, const float ext[12], float light,
                  const marl_sim_params* sim, float deltas[12]);
/* Returns 1 and writes the target when a division target exists; *draws is
 * advanced by the number of RNG draws consumed. */
int orc_find_empty_neighbor(const uint16_t pos[3], const int32_t* cell_at, const marl_dims* d,
                            const marl_sim_params* sim, uint64_t seed, uint64_t stream, uint64_t entity,
                            uint64_t* draws, uint16_t out[3]);
void orc_mutate(marl_ruleset* r, const marl_sim_params* sim, uint64_t seed, uint64_t stream,
                uint64_t entity, uint64_t* draws);
float orc_normal(uint64_t seed, uint64_t stream, uint64_t entity, uint64_t* draws, float sd);

/* hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27) for each pair k of
 * pairs = (donor, recipient) cell indices, in order; pair k draws from entity
 * k of stream ORC_TAG_HGT << 32 under key seed. */
void orc_hgt_transfer(marl_ruleset* rules, const int32_t* pairs, int64_t n_pairs, uint64_t seed);

/* swap_remove compaction of deaths (sorted ascending, distinct), sequential
 * emulation; writes the surviving original indices in final order to perm. */
int64_t orc_compact_indices(int64_t n, const int64_t* deaths, int64_t k, int64_t* perm);

typedef struct {
  int64_t divisions, deaths, births_added, n_cells;
} orc_tick_stats;

/* One full tick (crates/marl-sim/src/lib.rs:160-277).  Arrays are grown by the
 * caller: capacity must be >= n + number of divisions (n*2 is always enough). */
int orc_tick(const marl_dims* d, const marl_sim_params* sim, uint64_t tick, float* field, float* scratch,
             float* light, marl_cell_state* cells, marl_ruleset* rules, int64_t* n, int64_t capacity,
             uint8_t* events_out, orc_tick_stats* st, int nthreads);

/* Cell pass only (lib.rs:204-217) with the given field/light; events to events_out. */
void orc_cell_pass(const marl_dims* d, const marl_sim_params* sim, float* field, const float* light,
                   marl_cell_state* cells, const marl_ruleset* rules, int64_t n, uint8_t* events_out);

/* seed_cells (crates/marl-sim/src/seeding.rs:28-58) with the project's RNG;
 * returns the number of cells placed; positions/lineages to out arrays. */
int64_t orc_seed_positions(const marl_dims* d, uint64_t seed, uint64_t band, int64_t count, int32_t margin,
                           int32_t z_lo, int32_t z_hi, int32_t* cell_at, uint16_t* pos_out,
                           uint64_t* lineage_out, int64_t index_base);

#ifdef __cplusplus
}
#endif
#endif
