// pending.cpp -- entry points not yet implemented; they report
// ALLS_E_UNSUPPORTED instead of doing work.
#include "host_internal.h"

extern "C" {

alls_status alls_pattern_create(alls_ctx*, const alls_mesh*, const alls_opts*, alls_pattern** pat, alls_csr*) {
    if (pat) *pat = nullptr;
    return ALLS_E_UNSUPPORTED;
}
alls_status alls_assemble_values(alls_ctx*, const alls_pattern*, const double*, int32_t, alls_csr*) {
    return ALLS_E_UNSUPPORTED;
}
int64_t alls_pattern_nnz(const alls_pattern*) { return -1; }
void alls_pattern_free(alls_ctx*, alls_pattern*) {}
alls_status alls_color(alls_ctx*, const alls_pattern*, alls_coloring** col) {
    if (col) *col = nullptr;
    return ALLS_E_UNSUPPORTED;
}
alls_status alls_coloring_from_host(alls_ctx*, const alls_pattern*, const int32_t*, alls_coloring** col) {
    if (col) *col = nullptr;
    return ALLS_E_UNSUPPORTED;
}
int32_t alls_coloring_count(const alls_coloring*) { return -1; }
alls_status alls_coloring_download(alls_ctx*, const alls_coloring*, int32_t*) { return ALLS_E_UNSUPPORTED; }
void alls_coloring_free(alls_ctx*, alls_coloring*) {}
alls_status alls_assemble_colored(alls_ctx*, const alls_pattern*, const alls_coloring*, const double*, int32_t,
                                  alls_csr*) {
    return ALLS_E_UNSUPPORTED;
}
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
