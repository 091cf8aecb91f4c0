// Batched saddle searches (placeholder entry points; the persistent search kernel lands here).
#include "plain.hpp"

extern "C" {

void flyg_batch_default_options(flyg_batch_options* o) { *o = flyg_batch_options{FLYG_MIXED, 0, 0, 0.0, 0}; }

int flyg_batch_create(flyg_ctx ctx, flyg_eam, const flyg_dimer_options*, const flyg_rotor_options*,
                      const flyg_batch_options*, const double*, const int*, const double*, const uint32_t*, int64_t,
                      flyg_batch*) {
  ctx->err = "batched searches are not built in this library";
  return FLYG_INVALID_ARG;
}
int flyg_batch_destroy(flyg_batch) { return FLYG_OK; }
int flyg_batch_submit(flyg_batch, int, const int64_t*, const double*, const double*, const double*, int64_t*) {
  return FLYG_INVALID_ARG;
}
int flyg_batch_run(flyg_batch) { return FLYG_INVALID_ARG; }
int flyg_batch_result(flyg_batch, int64_t, flyg_search_stats*, double*, double*) { return FLYG_INVALID_ARG; }
int flyg_batch_region(flyg_batch, int64_t, int*, int*) { return FLYG_INVALID_ARG; }

}  // extern "C"
