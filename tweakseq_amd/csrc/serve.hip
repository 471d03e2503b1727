// Serving loop (continuous batching) -- placeholder until implemented.
#include <hip/hip_runtime.h>
#include "runtime_internal.h"

void serve_shutdown(ta_ctx *) {}
int ta_serve_config(ta_ctx *, int64_t, int32_t) { return TA_E_INVALID; }
int ta_submit(ta_ctx *, const ta_params *, const ta_pair *, int32_t, uint64_t *) { return TA_E_INVALID; }
int ta_poll(ta_ctx *, uint64_t, ta_result *, char *, int64_t, int64_t *) { return TA_E_INVALID; }
int ta_cancel(ta_ctx *, uint64_t) { return TA_E_INVALID; }
