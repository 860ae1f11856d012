// pending.cpp -- entry points not yet implemented; they report
// ALLS_E_UNSUPPORTED instead of doing work.
#include "host_internal.h"

extern "C" {

alls_status alls_assemble_batch(alls_ctx*, int32_t, const alls_mesh*, const alls_opts*, alls_csr*) {
    return ALLS_E_UNSUPPORTED;
}
alls_status alls_server_create(const alls_server_config*, alls_server** srv) {
    if (srv) *srv = nullptr;
    return ALLS_E_UNSUPPORTED;
}
alls_status alls_server_submit(alls_server*, const alls_mesh*, uint64_t, const alls_opts*, int64_t*) {
    return ALLS_E_UNSUPPORTED;
}
alls_status alls_server_wait(alls_server*, int64_t, alls_csr*, alls_request_stats*) { return ALLS_E_UNSUPPORTED; }
void alls_server_release(alls_server*, alls_csr*) {}
void alls_server_destroy(alls_server*) {}

}  // extern "C"
