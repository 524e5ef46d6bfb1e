// Entry points whose engines are not linked into this build yet: they report
// AB_INVALID_ARGUMENT with a message instead of being absent.
#include "context.hpp"

using namespace ab;

extern "C" {
ab_status ab_server_start(ab_context*, uint64_t) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_server_stop(ab_context*) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_job_submit(ab_context*, const ab_job_desc*, uint64_t*) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_job_poll(ab_context*, uint64_t, ab_job_status*, float*) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_job_wait(ab_context*, uint64_t, ab_job_status*, float*) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_job_cancel(ab_context*, uint64_t) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_job_release(ab_context*, uint64_t) { return fail(AB_INVALID_ARGUMENT, "not available"); }
ab_status ab_server_stats_get(ab_context*, ab_server_stats*) { return fail(AB_INVALID_ARGUMENT, "not available"); }
}
