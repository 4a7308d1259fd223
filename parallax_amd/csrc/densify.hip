// Adaptive density control on the device (placeholder until the compaction
// kernels land).
#include "../../include/splat_hip.h"
#include "splat_context.h"

extern "C" {

splat_status splat_densify(splat_context* ctx, const splat_densify_params*, float, int32_t, uint32_t*, int32_t*,
                           int32_t*) {
  if (ctx) ctx->error = "densify is not available in this build";
  return SPLAT_ERROR_STATE;
}

splat_status splat_reset_opacity(splat_context* ctx, float) {
  if (ctx) ctx->error = "reset_opacity is not available in this build";
  return SPLAT_ERROR_STATE;
}

splat_status splat_prune_and_sort(splat_context* ctx, float) {
  if (ctx) ctx->error = "prune_and_sort is not available in this build";
  return SPLAT_ERROR_STATE;
}

}  // extern "C"
