/*
 * marl_hip.h — C ABI of the MI355X (gfx950) per-tick field-and-cell path of MARL.
 *
 * Every entry point replaces a piece of the reference engine's tick
 * (alvertremantel/MARL-OCHRE, Rust).  The reference seam is the `gpu` Cargo
 * feature's `GpuFieldDiffuser` (crates/marl-gpu/src/field_diffusion.rs:31-268),
 * selected by `--gpu-diffusion` (crates/marl-engine/src/main.rs:3-6).  This ABI
 * keeps that contract for the drop-in call and adds a GPU-resident tick that
 * covers the rest of `marl_sim::run`'s loop body (crates/marl-sim/src/lib.rs:160-277).
 *
 * Conventions
 *   - every call returns an int status (MARL_OK or one of MARL_ERR_*);
 *     the codes mirror `GpuError` (crates/marl-gpu/src/context.rs:4-21);
 *   - on error host buffers are left untouched (the reference changes the host
 *     field only after a successful readback, field_diffusion.rs:285-290) and
 *     marl_last_error() returns a message;
 *   - host buffers are owned by the caller, device memory by the context;
 *   - one host thread per context; calls block until the GPU work is done
 *     unless the name says otherwise (the `_async` timing helpers);
 *   - for tests of a caller's fallback, the environment variable
 *     MARL_HIP_FAULT=init makes marl_create fail with MARL_ERR_DEVICE_INIT,
 *     and MARL_HIP_FAULT=tick:<n> makes the n-th marl_diffuse_tick or
 *     marl_tick call of each context fail with MARL_ERR_TRANSFER before it
 *     changes any state (read at every call).
 *
 * Layouts at the boundary
 *   - field AoS: ((z*Y + y)*X + x)*12 + s, f32 (crates/marl-field/src/field.rs:33-35);
 *   - field SoA: (s*Z + z)*Y*X + y*X + x, f32 (species planes; the device
 *     holds species quads, ((s/4)*Z*Y*X + (z*Y + y)*X + x)*4 + s%4);
 *   - occupancy: one byte per voxel, (z*Y + y)*X + x, nonzero = occupied
 *     (crates/marl-sim/src/lib.rs:174-179);
 *   - light: (z*Y + y)*X + x, f32 (crates/marl-field/src/light.rs:20-24);
 *   - cells: marl_cell_state[n] + marl_ruleset[n] in the engine's `cells` Vec order.
 */
#ifndef MARL_HIP_H
#define MARL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MARL_S_EXT 12          /* marl-config/src/lib.rs:22 */
#define MARL_M_INT 16          /* marl-config/src/lib.rs:23 */
#define MARL_R_MAX 16          /* marl-config/src/lib.rs:26 */
#define MARL_S_RECEPTORS 8     /* marl-config/src/lib.rs:27 */
#define MARL_S_TRANSPORTERS 8  /* marl-config/src/lib.rs:28 */
#define MARL_S_EFFECTORS 8     /* marl-config/src/lib.rs:29 */

/* Status codes (≙ GpuError, crates/marl-gpu/src/context.rs:4-9). */
enum {
  MARL_OK = 0,
  MARL_ERR_NO_DEVICE = 1,     /* ≙ GpuError::NoAdapter */
  MARL_ERR_DEVICE_INIT = 2,   /* ≙ GpuError::DeviceRequest */
  MARL_ERR_TRANSFER = 3,      /* ≙ GpuError::BufferMap (and any failed HIP call after init) */
  MARL_ERR_INVALID_INPUT = 4  /* ≙ GpuError::InvalidInput */
};

/* Cell events (≙ CellEvent, crates/marl-cell/src/cell.rs:154-160). */
enum { MARL_EVENT_NONE = 0, MARL_EVENT_DIVISION = 1, MARL_EVENT_DEATH = 2, MARL_EVENT_QUIESCENCE = 3 };

/* Kinetics modes for the resident tick. */
enum { MARL_KINETICS_PLAIN = 0, MARL_KINETICS_SURROGATE = 1, MARL_KINETICS_SHADOW = 2 };

/* Field layout of the plain cell pass (marl_set_cell_layout). */
enum { MARL_CELL_LAYOUT_AUTO = 0, MARL_CELL_LAYOUT_AOS = 1, MARL_CELL_LAYOUT_QUADS = 2 };

typedef struct marl_ctx marl_ctx;
typedef struct marl_ensemble marl_ensemble;

/* Grid size.  The reference fixes it at compile time (marl-config/src/lib.rs:17-19);
 * here it is per context.  x must be a multiple of 32. */
typedef struct {
  int32_t x, y, z;
} marl_dims;

/* The SimulationConfig fields the path reads (marl-config/src/lib.rs:41-106,
 * defaults :108-168), plus the seed of the counter-based RNG that replaces the
 * reference's OS-seeded ThreadRng (crates/marl-sim/src/lib.rs:35). */
typedef struct {
  float dt;
  int32_t diffusion_substeps;
  float d_voxel[MARL_S_EXT];
  float lambda_decay[MARL_S_EXT];
  float source_rate_oxidant, source_rate_carbon, source_rate_reductant;
  float epsilon, c_max, lambda_maintenance, hard_death_floor, reaction_maintenance;
  float base_division_prep, prep_maintenance_multiplier, rush_penalty_rate;
  float alpha_eps, k_eps;
  float light_efficiency, surface_intensity, cell_absorption, chemical_absorption, light_floor;
  float mutation_stddev, structural_mutation_rate_mult, meta_mutation_rate;
  float meta_mutation_clamp_low, meta_mutation_clamp_high;
  float hill_exponent_clamp_low, hill_exponent_clamp_high, active_reaction_threshold;
  int32_t division_neighbor_distance;
  uint32_t _pad0;
  uint64_t rng_seed;
} marl_sim_params;

/* Ruleset parameter records (crates/marl-cell/src/cell.rs:50-151). */
typedef struct { float k_half, n_hill, gain; } marl_receptor;                 /* 12 B */
typedef struct { float uptake_rate, secrete_rate; uint8_t ext_species, int_species, _pad[2]; } marl_transport; /* 12 B */
typedef struct { uint8_t substrate, product, catalyst, cofactor; float k_m, v_max, k_cat; } marl_reaction;   /* 16 B */
typedef struct { float threshold, rate; uint8_t int_species, ext_species, _pad[2]; } marl_effector;          /* 12 B */

typedef struct {
  marl_receptor receptors[MARL_S_RECEPTORS];      /*   0: 96 B */
  marl_transport transport[MARL_S_TRANSPORTERS];  /*  96: 96 B */
  marl_reaction reactions[MARL_R_MAX];            /* 192: 256 B */
  marl_effector effectors[MARL_S_EFFECTORS];      /* 448: 96 B */
  float division_energy, death_energy, quiescence_energy, division_prep_ticks; /* 544: FateParams */
  float hgt_propensity, mutation_rate;            /* 560 */
  float _reserved[2];                             /* 568 → 576 B */
} marl_ruleset;

/* Cell state (crates/marl-cell/src/cell.rs:168-183), 96 B. */
typedef struct {
  float internal[MARL_M_INT];   /*  0: 64 B */
  uint16_t pos[3];              /* 64 */
  uint8_t quiescent;            /* 70 */
  uint8_t starter_type;         /* 71 */
  uint64_t lineage_id;          /* 72 */
  uint32_t age;                 /* 80 */
  uint16_t prep_remaining;      /* 84 */
  uint16_t _pad;                /* 86 */
  uint32_t _reserved[2];        /* 88 → 96 B */
} marl_cell_state;

/* Per-tick counters of the resident tick. */
typedef struct {
  int64_t n_cells;        /* population after compaction */
  int64_t divisions;      /* ≙ tick_divisions (lib.rs:243) */
  int64_t deaths;         /* ≙ tick_deaths (lib.rs:248) */
  int64_t births_added;   /* births that survived the duplicate-position check (lib.rs:270-277) */
  int32_t levels;         /* depth of the conflict schedule of the cell pass */
  int32_t _pad;
  float ms_field;         /* occupancy rebuild, sources, substeps, light (hipEvents) */
  float ms_schedule;      /* empty-neighbour masks and the conflict levels (beside the field phase);
                             0 when a population of at most 384 cells runs its whole cell phase
                             in one workgroup */
  float ms_kinetics;      /* the pass's field view, then the per-level kinetics launches (with the
                             one-workgroup cell phase: all of it, schedule and fate included) */
  float ms_fate;          /* divisions, mutation, deaths, compaction (and the transpose back beside them) */
  float ms_tick;          /* the whole tick, first launch to last */
  float ms_transpose_back;/* the pass's field view back to the resident layout, beside fate */
} marl_tick_stats;

/* ---- context ---------------------------------------------------------- */

/* ≙ GpuFieldDiffuser::new (crates/marl-gpu/src/field_diffusion.rs:32-156).
 * Selects device 0, allocates the field double buffer, the occupancy bit mask
 * and the light plane.  Returns MARL_ERR_NO_DEVICE when no GPU is visible. */
int marl_create(const marl_dims* dims, marl_ctx** out);
/* ≙ Drop for GpuFieldDiffuser: frees every device buffer. */
void marl_destroy(marl_ctx* ctx);
/* Message of the last failed call (≙ GpuError's Display, context.rs:11-20). */
const char* marl_last_error(const marl_ctx* ctx);
/* Number of visible GPUs; 0 when none (no error). */
int marl_device_count(void);
/* Library version string. */
const char* marl_version(void);

/* ---- drop-in (same contract as the reference seam) ---------------------- */

/* ≙ GpuFieldDiffuser::diffuse_tick_with_cells (field_diffusion.rs:158-268):
 * uploads the AoS field and the byte occupancy, runs sim->diffusion_substeps
 * masked substeps with dt_sub = dt / substeps, and writes the result back into
 * field_aos.  Length checks as :164-175 (n_floats == X*Y*Z*12, n_voxels ==
 * X*Y*Z), no-op OK when substeps == 0 (:176-178). */
int marl_diffuse_tick(marl_ctx* ctx, float* field_aos, size_t n_floats,
                      const uint8_t* occupancy, size_t n_voxels, const marl_sim_params* sim);

/* ---- resident field ----------------------------------------------------- */

int marl_upload_field_aos(marl_ctx* ctx, const float* field_aos, size_t n_floats);
int marl_download_field_aos(marl_ctx* ctx, float* field_aos, size_t n_floats);
int marl_upload_field_soa(marl_ctx* ctx, const float* field_soa, size_t n_floats);
int marl_download_field_soa(marl_ctx* ctx, float* field_soa, size_t n_floats);
/* Byte occupancy → device bit mask (for field-only use; the resident tick
 * rebuilds the mask from the cell list, lib.rs:174-179). */
int marl_upload_occupancy(marl_ctx* ctx, const uint8_t* occupancy, size_t n_voxels);
int marl_download_occupancy(marl_ctx* ctx, uint8_t* occupancy, size_t n_voxels);

/* ≙ Field::apply_boundary_sources (crates/marl-field/src/field.rs:231-250). */
int marl_apply_boundary_sources(marl_ctx* ctx, const marl_sim_params* sim);
/* ≙ Field::diffuse_tick_with_cells (field.rs:212-217) on the resident field
 * with the resident occupancy: n_substeps substeps of dt_sub = dt / sim->diffusion_substeps. */
int marl_diffuse_substeps(marl_ctx* ctx, const marl_sim_params* sim, int32_t n_substeps);
/* ≙ LightField::update (crates/marl-field/src/light.rs:29-59). */
int marl_light_update(marl_ctx* ctx, const marl_sim_params* sim);
int marl_download_light(marl_ctx* ctx, float* light, size_t n_voxels);
int marl_upload_light(marl_ctx* ctx, const float* light, size_t n_voxels);

/* ---- resident cells and the full tick ------------------------------------ */

/* Upload the population in `cells` Vec order (crates/marl-sim/src/lib.rs:56-57).
 * Positions must be distinct and in bounds. */
int marl_upload_cells(marl_ctx* ctx, const marl_cell_state* cells, const marl_ruleset* rulesets, size_t n);
int marl_cell_count(marl_ctx* ctx, int64_t* n);
int marl_download_cells(marl_ctx* ctx, marl_cell_state* cells, marl_ruleset* rulesets, size_t capacity, int64_t* n_out);
/* One or more full ticks (lib.rs:164-277): boundary sources, occupancy rebuild,
 * diffusion, light, the ordered cell pass, fate (division with mutation,
 * death) and compaction.  `tick0` is the tick number of the first tick (keys
 * the RNG).  Stats of the last tick go to *stats when non-null. */
int marl_tick(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick0, int32_t n_ticks, marl_tick_stats* stats);
/* Only the cell pass (lib.rs:204-217) on the resident field/light/cells;
 * events (MARL_EVENT_*) per cell to `events` when non-null (length >= n). */
/* The same tick as marl_tick, captured once as a hipGraph and replayed: every
 * count (cells, schedule depth, births, deaths) lives on the device, so a
 * tick costs one graph launch and one small readback instead of a host round
 * trip per phase.  Results are bit-identical to marl_tick.  Requires plain
 * kinetics and no harvest; an odd diffusion_substeps alternates between two
 * graphs, one per field-buffer parity.  A tick whose schedule needs more
 * relaxation sweeps or levels than the captured budget (see
 * marl_set_graph_limits) is finished by the host-driven path; *host_finished
 * (may be NULL) counts such ticks.  stats->ms_field holds the graph's GPU
 * time; the other phase times are 0. */
int marl_tick_graph(marl_ctx* ctx, const marl_sim_params* sim, int64_t tick0, int32_t n_ticks,
                    marl_tick_stats* stats, int32_t* host_finished);
/* Field layout the plain cell pass works on; results are identical either way.
 * MARL_CELL_LAYOUT_AOS: an AoS view (48-byte voxel records) that a tick's
 * last substep writes (a pass outside a tick transposes it from the quads),
 * transposed back to the resident species quads after the pass; _QUADS: the
 * resident quads in place (no transposes, more cache lines per neighbour);
 * _AUTO (default): quads below about one cell per 20 voxels (per 10 for
 * ensemble columns, which transpose both ways).  A tuning knob
 * with no counterpart in the reference (whose field is always AoS,
 * crates/marl-field/src/field.rs:15-35). */
int marl_set_cell_layout(marl_ctx* ctx, int32_t layout);
/* Threads per block of the plain cell pass: 64, 128 (default) or 256; results
 * are identical for every size.  A tuning knob with no reference counterpart. */
int marl_set_kinetics_block(marl_ctx* ctx, int32_t block);
int marl_get_kinetics_block(marl_ctx* ctx, int32_t* block);
/* The project's kinetics autotuner: times the plain cell pass (the ordered
 * per-level launches with the field layout's transposes) on the resident
 * field, light and population with each block size, keeps the fastest and
 * writes it to *block_out.  Field and cells are saved first and restored after
 * every trial, so the resident state is unchanged. */
int marl_autotune_kinetics(marl_ctx* ctx, const marl_sim_params* sim, int32_t* block_out);
/* Budget of relaxation sweeps and level launches in the captured tick (1..64
 * each; a change captures the graph again).  (0, 0) selects the default
 * automatic budget: it starts at 64 each and then follows what the schedule
 * needed, with a slack of need / 2 + 2 (3 to 8): within 1 of the need it
 * grows to the need + 2 slacks, more than 3 slacks above it shrinks to the
 * need + 1 slack. */
int marl_set_graph_limits(marl_ctx* ctx, int32_t sweeps, int32_t levels);
int marl_cell_pass(marl_ctx* ctx, const marl_sim_params* sim, uint8_t* events, size_t n_events, int32_t* levels_out);
/* FP64 per-species totals: out[0..11] empty voxels, out[12..23] occupied
 * voxels, out[24..39] cell internals (no reference counterpart: the reference
 * keeps no mass budget, docs/INFO.md:276). */
int marl_mass_budget(marl_ctx* ctx, double* out, size_t n_out);
/* Per-phase mass budget of the host-driven tick (no reference counterpart):
 * with it on, every marl_tick records marl_mass_budget's 40 totals at four
 * points of the tick — after the occupancy rebuild (before the sources), after
 * the field phase (sources, substeps, light), after the cell pass, and after
 * fate and compaction (against the new occupancy) — so that the matter each
 * phase adds or removes can be read off.  Costs four FP64 reductions a tick. */
int marl_set_phase_budget(marl_ctx* ctx, int32_t on);
/* The last such tick's budgets, out[phase*40 + k] (n_out >= 160). */
int marl_tick_phase_budget(marl_ctx* ctx, double* out, size_t n_out);

/* ---- setup ----------------------------------------------------------------- */

/* ≙ seed_cells (crates/marl-sim/src/seeding.rs:28-58): up to count*8 attempts
 * drawing x in [margin, X-margin), y in [margin, Y-margin), z in [z_lo,
 * min(z_hi, Z)), skipping taken voxels, then a u64 lineage per placed cell —
 * from the Philox stream (seed, 1<<32 | band).  `cell_at` (host, X*Y*Z int32,
 * -1 = free) is read and updated; placed cells get indices index_base + k.
 * Host-only: needs no device. */
int marl_seed_positions(const marl_dims* dims, uint64_t seed, uint64_t band, int64_t count, int32_t margin,
                        int32_t z_lo, int32_t z_hi, int32_t* cell_at, uint16_t* pos_out, uint64_t* lineage_out,
                        int64_t index_base, int64_t* placed);

/* ≙ hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27), which the
 * reference's run() never calls: for each pair k of `pairs` (n_pairs x
 * {donor, recipient} cell indices into the resident population) a donor
 * reaction with |v_max| > 1e-9, drawn uniformly, overwrites the recipient's
 * first reaction with |v_max| < 1e-9, or a drawn slot when it has none.  Pair k
 * draws from the Philox stream (seed, 3<<32), entity k: draw 0 picks the donor
 * reaction, draw 1 the slot.  The pairs are independent transfers: recipients
 * must be distinct and no recipient may be a donor (else INVALID_INPUT, state
 * untouched). */
int marl_hgt_transfer(marl_ctx* ctx, const int32_t* pairs, size_t n_pairs, uint64_t seed);

/* ---- timing helpers (hipEvents on the context's stream) ------------------ */

/* Launch n substeps back to back on the context stream and return the elapsed
 * GPU time between hipEvents recorded around them. */
int marl_time_diffuse(marl_ctx* ctx, const marl_sim_params* sim, int32_t n_substeps, float* ms_out);
int marl_time_light(marl_ctx* ctx, const marl_sim_params* sim, int32_t n_updates, float* ms_out);
/* n_ticks field phases as a tick runs them (boundary sources when
 * with_sources, diffusion_substeps substeps, the light scan; sources fused
 * into the first substep and, when the tile's z-chunk spans the column, the
 * light scan into the last), timed like marl_time_diffuse. */
int marl_time_field_phase(marl_ctx* ctx, const marl_sim_params* sim, int32_t n_ticks, int32_t with_sources,
                          float* ms_out);
/* Choose stencil tile shape by timing the candidates on the resident field
 * (the project's autotuner).  Writes the chosen tile id. */
int marl_autotune_stencil(marl_ctx* ctx, const marl_sim_params* sim, int32_t* tile_out);
int marl_set_stencil_tile(marl_ctx* ctx, int32_t tile);
int marl_get_stencil_tile(marl_ctx* ctx, int32_t* tile);
int marl_stencil_tile_count(void);
/* Substeps one launch of a tile id performs: 1, or 2 for the temporally
 * blocked tiles (which keep the intermediate substep on chip). */
int marl_stencil_tile_substeps(int32_t tile);
/* Rows per thread, z-chunk and prefetch distance (planes) of a tile id. */
int marl_stencil_tile_shape(int32_t tile, int32_t* rows_per_thread, int32_t* z_chunk, int32_t* prefetch);
/* Species per thread and waves per block of a tile id: a block of W waves
 * holds W/4 groups of 4 waves sharing one LDS plane, each taking that many
 * species. */
int marl_stencil_tile_split(int32_t tile, int32_t* species_per_thread, int32_t* waves_per_block);
/* Plain copy of the resident field into the scratch buffer (n_reps times,
 * float4 loads/stores): the HBM calibration kernel for roofline and counter
 * scaling.  *bytes_out = bytes read + written per rep. */
int marl_copy_probe(marl_ctx* ctx, int32_t n_reps, float* ms_out, double* bytes_out);
/* Random-access calibration (no reference counterpart): n_reps launches in
 * which every 4 lanes read one 64-byte sector at a hashed position of a 1 GiB
 * buffer the call allocates and frees, 4 sectors per thread; the ceiling of
 * kernels that read scattered records.  *bytes_out = sector bytes read per rep. */
int marl_gather_probe(marl_ctx* ctx, int32_t n_reps, float* ms_out, double* bytes_out);
/* Dense BF16 MFMA calibration (no reference counterpart): n_reps launches in
 * which every wave of 4 resident 256-thread blocks per CU runs 4 independent
 * chains of 4,096 v_mfma_f32_32x32x16_bf16 on register operands; the measured
 * denominator beside the vendor's 2.5 PFLOP/s.  *flop_out = FLOP per rep. */
int marl_mfma_probe(marl_ctx* ctx, int32_t n_reps, float* ms_out, double* flop_out);
int marl_sync(marl_ctx* ctx);

/* ---- outputs: reductions and packing on the resident state ---------------- */

/* The population figures of DataLogger::log_tick (crates/marl-output/src/data.rs:195-275)
 * and print_stats (crates/marl-sim/src/stats.rs:9-82): FP64 sums of internal[0]
 * (energy), [5] and [6] (enzymes a/b), the number of reactions with
 * |v_max| > 1e-9, and cells with z < Z/3, < 2*(Z/3) and the rest (Z/3 in
 * integers, stats.rs:39).  The FP64 sums are deterministic but summed in a
 * different order from the reference's sequential loop. */
typedef struct {
  int64_t n_cells;
  double sum_energy, sum_enzyme_a, sum_enzyme_b;
  int64_t sum_active_reactions;
  int64_t z_thirds[3];
} marl_population_summary;
/* z_counts (may be NULL): cells per z-layer, nz >= Z. */
int marl_cell_summary(marl_ctx* ctx, marl_population_summary* out, int64_t* z_counts, size_t nz);
/* marl_cell_summary without allocation.  No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results, bit for bit, as marl_cell_summary, which allocates and
 * frees its partial sums on every call; this variant keeps them in a fixed
 * 72 KiB device array per GPU, set up by marl_create, and calls no allocator.
 * Grids deeper than 4096 layers are rejected (MARL_ERR_INVALID_INPUT).  Every
 * context on one GPU shares that array: calls that overlap on one GPU (from
 * several host threads) race. */
int marl_cell_summary_noalloc(marl_ctx* ctx, marl_population_summary* out, int64_t* z_counts, size_t nz);
/* ≙ DataLogger::snapshot_chemistry (data.rs:292-335): per z-layer, the FP64 mean
 * over the XY plane of each of the 12 species, then of light: out[z*13 + c]. */
int marl_chemistry_profile(marl_ctx* ctx, double* out, size_t n);
/* ≙ write_cell_dump's records (crates/marl-output/src/binary_dump.rs:34-41,
 * crates/marl-format/src/lib.rs:350-364): 25 bytes per cell, packed on the GPU,
 * {pos: f32[3], lineage_id: u64, starter_type: u8, energy: f32}, little-endian. */
int marl_download_viewer_cells(marl_ctx* ctx, uint8_t* out, size_t cap_bytes, int64_t* n_out);
/* Point reads (≙ Field::get / LightField::get at stats.rs:51-56, 86-103):
 * queries[4k..4k+3] = x, y, z, species (-1 = light). */
int marl_sample(marl_ctx* ctx, const int32_t* queries, size_t n, float* out);

/* ---- ensembles ------------------------------------------------------------- */
/* No reference counterpart: the reference runs one column per process
 * (crates/marl-sim/src/lib.rs:160-277).  An ensemble advances independent
 * columns of uneven sizes together: each column is an ordinary context
 * (created, uploaded and downloaded as usual; not owned by the ensemble).
 * A tick is one hipGraph: every column's field phase in one persistent
 * stencil launch per substep and one light launch, and the columns' cell
 * phases side by side on their own streams.  Column c uses seeds[c] as its
 * rng_seed (sim->rng_seed + c when seeds is NULL); results are bit-identical
 * to marl_tick on each column alone.  Needs an even, positive substep count,
 * plain kinetics and no harvest. */
int marl_ensemble_create(marl_ctx* const* columns, int32_t n_columns, marl_ensemble** out);
void marl_ensemble_destroy(marl_ensemble* e);
const char* marl_ensemble_last_error(const marl_ensemble* e);
/* stats: n_columns entries (may be NULL); *host_finished counts column-ticks
 * whose cell phase was finished host-driven; *ms_out = GPU time of the graphs. */
int marl_ensemble_tick(marl_ensemble* e, const marl_sim_params* sim, const uint64_t* seeds, int64_t tick0,
                       int32_t n_ticks, marl_tick_stats* stats, int32_t* host_finished, float* ms_out);
int marl_ensemble_set_tile(marl_ensemble* e, int32_t tile);
/* Grid of the batched stencil launch: 0 = persistent (as many blocks as fit
 * the chip, each walking its share of the tiles), 1 = one block per tile,
 * 2 = persistent over balanced pieces (a multiple of 8 blocks; per XCD a run
 * of neighbouring tiles of equal work, each tile marching its column's whole
 * depth except in the last, partial round, whose tiles are cut into z-chunks;
 * the z-chunk settings then do not apply and the light scan is its own
 * launch).  Results are identical in every mode. */
int marl_ensemble_set_grid(marl_ensemble* e, int32_t per_tile);
int marl_ensemble_get_grid(const marl_ensemble* e, int32_t* mode);
/* Average time of one batched substep over all columns (fields unchanged). */
int marl_ensemble_time_field(marl_ensemble* e, const marl_sim_params* sim, int32_t n_substeps, float* ms_out);
/* Average time of a tick's batched field phase (the substeps, sources in the
 * first, and the light scan) over n_ticks, reading the current buffers each
 * time: the fields are unchanged, the light planes are rewritten. */
int marl_ensemble_time_field_phase(marl_ensemble* e, const marl_sim_params* sim, int32_t n_ticks, float* ms_out);
/* Per-column z-chunks: 0 = the tile shape's chunk (cut at each column's
 * depth) with the light scan as its own batched launch; 1 = every column one
 * chunk deep, which lets the last substep of the tick carry each column's
 * light scan (no separate light launch).  Results are identical. */
int marl_ensemble_set_whole_columns(marl_ensemble* e, int32_t whole);
/* Per-column z-chunk lengths when not whole columns: chunks[c] planes for
 * column c (0 = the tile shape's chunk), cut at the column's depth; n = the
 * column count, or 0 to clear.  Results are identical for every choice. */
int marl_ensemble_set_chunks(marl_ensemble* e, const int32_t* chunks, int32_t n);
/* The z-chunk each column's batched stencil tiles use in grid modes 0 and 1
 * (n = the column count). */
int marl_ensemble_get_chunks(const marl_ensemble* e, int32_t* chunks, int32_t n);
/* Times a tick's batched field phase (substeps and light) for every stencil
 * tile shape, with a persistent grid (both z-chunk modes), one block per tile
 * (both z-chunk modes) and balanced pieces, on the resident columns; then, for the five fastest
 * shapes without whole columns, the z-chunk of each group of equally deep
 * columns (4, 8, 16, 24, 32, 48, 64 planes or the whole depth, deepest group
 * first, two passes); keeps the fastest.  Then each column's plain cell-pass block size
 * (marl_autotune_kinetics on its own population).  Fields, light planes and
 * cells are left unchanged. */
int marl_ensemble_autotune(marl_ensemble* e, const marl_sim_params* sim, int32_t* tile_out);
/* The batched stencil launch: tile shape, grid blocks, logical tiles (the
 * pieces in grid mode 2) and whether the light scan rides the last substep
 * (each may be NULL). */
int marl_ensemble_launch_shape(const marl_ensemble* e, int32_t* tile, int32_t* grid_blocks, int32_t* stencil_tiles,
                               int32_t* whole_columns);

/* ---- learned kinetics surrogate ------------------------------------------- */
/* No reference counterpart: the reference integrates every cell's kinetics
 * exactly (crates/marl-cell/src/cell.rs:223-329).  The surrogate is an MLP
 * (80 -> 256 -> 256 -> 256 -> 28, SiLU) trained on the GPU from pairs the plain
 * kinetics pass harvests.  It predicts the residual between the exact update
 * and one explicit step of the same rate laws from the pre-tick state, routed
 * through the cell's own wiring; the explicit step and its per-reaction flux
 * ratios are among the inputs, so a cell of any wiring is eligible.  A cell
 * uses the prediction when its standardised features lie within zmax and its
 * predicted energy is at least energy_margin from every fate threshold; all
 * other cells run the plain path.
 * Features (80): internal[0..14], quiescent, environment[12], light, in-prep,
 * explicit internal deltas [0..14], explicit exchange fluxes [12], each
 * reaction's explicit flux / (its substrate + 1e-6) [16], 7 zeros.
 * Targets (28): exact minus explicit internal deltas [16] (slot 15 is 0: the
 * tick sets internal 15 to the light) and exchange fluxes [12]. */

typedef struct {
  int64_t steps;       /* Adam steps */
  int32_t batch;       /* minibatch rows */
  int32_t reinit;      /* 1: fresh He-uniform initialisation even if trained */
  float lr, lr_final;  /* cosine schedule from lr to lr_final */
  float zmax;          /* out-of-distribution bound (<= 0: 6) */
  float energy_margin; /* fate margin on the predicted energy (< 0: 0.05) */
  uint64_t seed;
  int64_t max_pairs;   /* train on the first max_pairs harvested pairs (<= 0: all) */
} marl_train_config;

typedef struct {
  float loss_first, loss_last, val_loss; /* MSE on standardised targets */
  float ms;                              /* wall time of the call */
  int64_t pairs, pairs_train;
} marl_train_report;

typedef struct {
  int64_t cells;              /* cells that reached the surrogate pass */
  int64_t surrogate;          /* advanced by the surrogate */
  int64_t fallback_topology;  /* always 0: every wiring is eligible */
  int64_t fallback_ood;       /* feature outside zmax */
  int64_t fallback_margin;    /* predicted energy near a fate threshold */
  int64_t shadow;             /* cells run both ways and compared */
  int64_t shadow_within_1e3;  /* max relative internal error <= 1e-3 */
  int64_t shadow_fate_agree;  /* same fate event and fate state */
  float shadow_max_rel_err;
  int32_t _pad;
  int64_t err_hist[24];       /* log10 max relative error, half-decade bins from 1e-9 */
  int64_t shadow_within_1e3_species[16]; /* shadow cells whose internal k is within 1e-3 (k < 15) */
  int64_t fate_exact[4];      /* shadow cells' fate events (MARL_EVENT_*) from the exact update */
  int64_t fate_surrogate[4];  /* ... and from the prediction */
} marl_surr_stats;

/* MARL_KINETICS_PLAIN (exact), _SURROGATE (learned, with fallback) or
 * _SHADOW (exact, with every eligible cell also predicted and compared). */
int marl_set_kinetics_mode(marl_ctx* ctx, int32_t mode);
/* Start appending (features, targets, topology hash) of a sample_rate share
 * of cells from each plain kinetics pass, up to capacity pairs. */
int marl_harvest_begin(marl_ctx* ctx, int64_t capacity, float sample_rate);
int marl_harvest_end(marl_ctx* ctx, int64_t* n_pairs);
/* feat [n][80], targ [n][28] (internal deltas 16, exchange fluxes 12), hash [n]. */
int marl_harvest_download(marl_ctx* ctx, float* feat, float* targ, uint64_t* hash, size_t cap, int64_t* n_out);
int marl_harvest_upload(marl_ctx* ctx, const float* feat, const float* targ, const uint64_t* hash, size_t n);
int marl_surrogate_train(marl_ctx* ctx, const marl_train_config* cfg, marl_train_report* report);
int marl_surrogate_param_count(void);
/* Flat FP32 parameters: per layer W [in x out] row-major then b [out];
 * norm = feature mean[80], std[80], target mean[28], std[28]. */
int marl_surrogate_get_params(marl_ctx* ctx, float* params, size_t n, float* norm, size_t n_norm);
int marl_surrogate_set_params(marl_ctx* ctx, const float* params, size_t n, const float* norm, size_t n_norm);
/* Network output for standardised inputs x [rows][80] -> y [rows][28]:
 * fused = 1 the inference kernel, 0 the training forward. */
int marl_surrogate_forward(marl_ctx* ctx, const float* x, size_t rows, float* y, int32_t fused);
/* MSE loss and its gradient (flat, as the parameters) for x, targets t. */
int marl_surrogate_grad(marl_ctx* ctx, const float* x, const float* t, size_t rows, float* loss, float* grads);
/* One Adam step (b1 0.9, b2 0.999, eps 1e-8) with the given gradient. */
int marl_surrogate_adam_step(marl_ctx* ctx, const float* grads, float lr);
/* zmax, energy margin, share of surrogate cells also run exactly for monitoring. */
int marl_surrogate_configure(marl_ctx* ctx, float zmax, float energy_margin, float shadow_rate);
/* Corrections whose magnitude in target-scale units (residual over the
 * output's gross explicit activity) is at most dead_zone are dropped, so the
 * explicit step stands where the network is not confident (default 0.2). */
int marl_surrogate_set_dead_zone(marl_ctx* ctx, float dead_zone);
int marl_surrogate_stats(marl_ctx* ctx, marl_surr_stats* out);
/* Average time of the fused inference kernel over `rows` synthetic rows
 * (hipEvents around `reps` back-to-back launches on the context stream). */
int marl_surrogate_time_infer(marl_ctx* ctx, int64_t rows, int32_t reps, float* ms_out);
int marl_surrogate_reset_stats(marl_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MARL_HIP_H */
