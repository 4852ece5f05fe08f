/* entry points of the whole-path stage that are not wired yet */
#include "msgs_capi_impl.h"
extern "C" {
int msgs_seed_batch(msgs_ctx* c, const msgs_reads*, const msgs_params*, int64_t*, int64_t*, double*, int64_t, int64_t*) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_map_batch(msgs_ctx* c, const msgs_reads*, const msgs_reads*, const msgs_params*, char*, int64_t, int64_t*, int64_t*) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_align_batch(msgs_ctx* c, const msgs_reads*, const msgs_params*, int64_t*, int64_t*, double*, char*, int64_t, int64_t, int64_t*) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_batch_upload(msgs_ctx* c, int32_t, const msgs_reads*, const msgs_reads*) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_batch_run(msgs_ctx* c, int32_t, const msgs_params*, int32_t) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_batch_stats(msgs_ctx* c, double*, int32_t) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
int msgs_batch_sam(msgs_ctx* c, char*, int64_t, int64_t*) { return fail(c, MSGS_E_UNSUPPORTED, "not yet"); }
}
