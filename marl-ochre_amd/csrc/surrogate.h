// surrogate.h — the learned kinetics surrogate's host-side state.
#pragma once

#include "kinetics.h"
#include "mlp.h"

constexpr int kErrBins = 24;  // log10 relative error histogram, bins of 0.5 decade from 1e-9

struct SurrogateState {
  MlpState mlp;
  // harvested pairs
  float* hfeat = nullptr;
  float* htarg = nullptr;
  uint64_t* hhash = nullptr;
  unsigned long long* hcount = nullptr;
  int64_t hcap = 0;
  uint64_t hthreshold = 0;
  bool harvesting = false;
  // model
  bool trained = false;
  float zmax = 6.0f;           // out-of-distribution bound on |standardised feature|
  float margin = 0.05f;        // fate margin on the predicted energy
  float dead_zone = 0.2f;      // corrections up to this (in target-scale units) are dropped
  uint64_t shadow_threshold = 0;  // sampled surrogate cells also run the plain path for monitoring
  float* inv = nullptr;        // 1 / feature std (device, kFeat), kept with the normalisation
  // per-level scratch
  uint16_t* featb = nullptr;
  float* out = nullptr;
  float* expl = nullptr;       // per cell of the level: explicit-step deltas [16] and fluxes [12], then
                               // their target scales [28] (target_scales)
  uint8_t* elig = nullptr;
  int64_t scratch_cap = 0;
  // statistics (device): counters as u64, error histogram, max error bits
  unsigned long long* stats = nullptr;
};

HarvestSink surrogate_harvest_sink(marl_ctx* ctx, const CellParams& p);
int surrogate_level(marl_ctx* ctx, const CellParams& p, const int32_t* order, int cnt, marl_cell_state* cells,
                    const marl_ruleset* rules, float* field);
void surrogate_free(SurrogateState* s);
