// noalloc.hip -- the fixed device scratch of alls_assemble_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the A-LLS
// pipeline with its list pool and consolidation arrays in one file-scope
// device array, `scratch`, of ALLS_NOALLOC_SCRATCH_BYTES.  Every GPU the
// library runs on holds its own copy until the process ends.  In host code
// the array's name is only a host-side stand-in: its device address is
// looked up once per device with hipGetSymbolAddress (at context creation,
// with the device current) and that pointer is what the kernels get.
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents; the pipeline clears what must start at zero (list headers and
// flags, with fill kernels).  Calls that overlap on one GPU race.
#include <mutex>

#include "host_internal.h"

namespace {

constexpr size_t ALLS_NOALLOC_SCRATCH_BYTES = size_t(128) << 20;   // 128 MiB per GPU
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[ALLS_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
bool g_done[MAX_DEVICES] = {};
hipError_t g_err[MAX_DEVICES] = {};

}  // namespace

namespace alls {

size_t noalloc_scratch_bytes() { return ALLS_NOALLOC_SCRATCH_BYTES; }

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return g_err[device];
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_done[device] = true;
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    return e;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_done[device] ? g_ptr[device] : nullptr;
}

}  // namespace alls
