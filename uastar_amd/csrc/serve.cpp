// serve.cpp -- continuous-batching request loop (see ua_serve_* in include/uastar.h).
#include <hip/hip_runtime.h>
#include "../../include/uastar.h"

struct ua_ctx;
void ua_serve_release(ua_ctx *) {}
extern "C" int ua_serve_submit(ua_ctx *, const ua_query *, uint64_t *) { return UA_ERR_INVALID; }
extern "C" int ua_serve_step(ua_ctx *, int64_t *, int64_t *) { return UA_ERR_INVALID; }
extern "C" int ua_serve_poll(ua_ctx *, uint64_t *, ua_result *, int64_t, int64_t *) { return UA_ERR_INVALID; }
extern "C" int ua_serve_drain(ua_ctx *) { return UA_ERR_INVALID; }
