// ts_batch.cpp -- batching and serving entry points (filled in by the batched engine).
#include <string>

#include "ts_internal.hpp"

struct ts_batch {
    std::string err = "batching is not available in this build";
};

extern "C" {
int ts_batch_create(const ts_batch_config*, ts_batch** out) {
    if (out) *out = nullptr;
    return TS_ERR_UNSUPPORTED;
}
int ts_batch_destroy(ts_batch* b) {
    delete b;
    return TS_OK;
}
const char* ts_batch_last_error(const ts_batch* b) { return b ? b->err.c_str() : "null batch"; }
int ts_batch_submit(ts_batch*, const ts_config*, const double*, int, double, long long*) { return TS_ERR_UNSUPPORTED; }
int ts_batch_run(ts_batch*, int, int*) { return TS_ERR_UNSUPPORTED; }
int ts_batch_poll(ts_batch*, long long, int*, double*, ts_step_info*) { return TS_ERR_UNSUPPORTED; }
int ts_batch_start_server(ts_batch*) { return TS_ERR_UNSUPPORTED; }
int ts_batch_stop_server(ts_batch*) { return TS_ERR_UNSUPPORTED; }
}
