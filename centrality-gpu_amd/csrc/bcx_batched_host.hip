// Batched multi-source path (host driver).
#include "bcx_runtime.h"

struct BatchedState {};

int batched_component(bcx_ctx *ctx, bcx_dgraph *, bcx::DComp &, const bcx::Plan &, const bcx_opts *,
                      double *, bcx_stats *) {
    return ctx_fail(ctx, BCX_E_ARG, "batched path not available");
}
int batched_debug_source(bcx_ctx *ctx, const bcx_component *, int32_t, const bcx_opts *, int32_t *,
                         uint64_t *, double *, double *, int32_t *) {
    return ctx_fail(ctx, BCX_E_ARG, "batched path not available");
}
void batched_release(bcx_ctx *) {}
void batched_graph_release(bcx_ctx *, bcx_dgraph *) {}
