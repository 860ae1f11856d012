// pending.cpp -- entry points not yet implemented; they report
// ALLS_E_UNSUPPORTED instead of doing work.
#include "host_internal.h"

extern "C" {

alls_status alls_assemble_batch(alls_ctx*, int32_t, const alls_mesh*, const alls_opts*, alls_csr*) {
    return ALLS_E_UNSUPPORTED;
}
}  // extern "C"
