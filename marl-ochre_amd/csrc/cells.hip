// cells.hip — resident cell path (placeholder entry points until the kernels land).
#include "marl_internal.h"

extern "C" {
int marl_upload_cells(marl_ctx* ctx, const marl_cell_state*, const marl_ruleset*, size_t) {
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
int marl_cell_count(marl_ctx* ctx, int64_t* n) {
  if (n) *n = 0;
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
int marl_download_cells(marl_ctx* ctx, marl_cell_state*, marl_ruleset*, size_t, int64_t*) {
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
int marl_tick(marl_ctx* ctx, const marl_sim_params*, int64_t, int32_t, marl_tick_stats*) {
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
int marl_cell_pass(marl_ctx* ctx, const marl_sim_params*, uint8_t*, size_t, int32_t*) {
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
int marl_mass_budget(marl_ctx* ctx, double*, size_t) {
  return marl_fail(ctx, MARL_ERR_INVALID_INPUT, "cell path not built");
}
}
