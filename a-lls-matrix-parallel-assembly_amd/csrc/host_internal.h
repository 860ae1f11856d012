// host_internal.h -- host-side (C++) internals of liballs.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>

#include "alls.h"
#include "alls_internal.h"

namespace alls {

// Caching device allocator (free lists; no device-wide sync on free).
hipError_t dev_alloc(void** out, size_t bytes);
// While on, dev_alloc on this thread fails (hipErrorOutOfMemory) instead of
// allocating: the alls_assemble_noalloc guarantee.
void no_alloc_scope(bool on);
// The fixed device scratch of alls_assemble_noalloc (noalloc.hip): its
// address on `device`, looked up once per device (call with the device
// current), and its size.
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
void dev_free(void* p);
void dev_trim();
// Device arena: one reservation carved up by the threads bound to it
// (dev_alloc on such a thread takes from the arena first; dev_free of an
// arena block from any thread returns it to the arena).
struct Arena;
Arena* arena_create(size_t bytes);   // nullptr when the reservation fails
void arena_destroy(Arena* a);        // every block must have been freed
void arena_bind(Arena* a);           // this thread's arena (nullptr: none)

// Grow-only device buffer (owns its allocation; movable, not copyable).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    hipError_t ensure(size_t need);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// What the pipeline needs to know about a mesh before launching.
struct MeshInfo {
    int kind = 0;              // 0 Q4, 1 T3, 2 mixed Q4/T3, 3 Hex20
    int dpn = 2;               // DOF per node (= dim)
    int npe_max = 4;           // nodes per element (max)
    int64_t entries_bound = 0; // upper bound of list entries (sum of npe^2)
    bool f32 = false;          // FP32-MFMA Hex20 integration
};

// An instantiated hipGraph and the key of the arguments it was captured with.
struct GraphSlot {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    void reset();
};

// Per-stream scratch: list headers, pool, consolidation arrays and a device
// copy of host meshes.  A context owns one; the server owns one per slot.
struct Workspace {
    DevBuf hdr, keys, geid, vals, misc, u, big, tiles;
    DevBuf coords, conn, types, E;
    DevBuf bdesc;                   // batch: source-mesh table, boundary indices and offsets
    DevBuf bsplit;                  // batch split: per-request output table (server)
    int64_t* host_misc = nullptr;   // pinned readback: misc ints + scan total
    void* pinned = nullptr;         // pinned staging of host inputs (grow-only)
    size_t pinned_bytes = 0;
    // Copies host arrays to device through the pinned buffer (one async copy
    // each); pageable copies would block behind other streams' work.
    hipError_t upload(void* const* dst, const void* const* src, const size_t* bytes, int n, hipStream_t s);
    // The pinned buffer, grown to `bytes`, for callers that fill it directly.
    hipError_t staging(size_t bytes, char** host);
    hipEvent_t ev[8] = {};
    bool events = false;
    GraphSlot graph;                // the captured pipeline (batch path)
    // Fixed scratch (alls_assemble_noalloc): when set, the list pool and
    // consolidation arrays are carved from it, the fused pass is not used
    // and the output CSR is the caller's (checked, never reallocated).
    char* ext = nullptr;
    size_t ext_bytes = 0;
    hipError_t init();
    void destroy();
    size_t device_bytes() const;
    void release_device();          // frees every device buffer, keeps events and pinned memory
};

alls_status classify_mesh(const alls_mesh* m, const alls_opts& o, MeshInfo* mi, std::string* err);
// List chunk size (entries) used when the caller leaves chunk_slots at 0.
int default_chunk(int kind);
// Chunks of the list pool: one home chunk per node plus the overflow
// reserve (full: every chunk the entries could need).
int64_t pool_chunks(int64_t n_nodes, int64_t entries_bound, int C, int kind, bool full);
// Device view of the mesh; host arrays are copied into the workspace.
alls_status stage_mesh(Workspace& ws, const alls_mesh* m, const MeshInfo& mi, DevMesh* dm, hipStream_t s,
                       std::string* err);
// Full A-LLS pipeline on `s`: insert, count, scan, write.  Synchronises `s`.
// use_graph: the whole pipeline runs as ONE hipGraph into an output sized by
// the entries bound, captured on first use and replayed while its arguments
// (buffer addresses, sizes, material) stay the same; the flags and the total
// are read back once, after it.
alls_status run_alls(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o, alls_csr* out,
                     hipStream_t s, alls_stats* st, std::string* err, bool use_graph = false);
// Makes sure `out` can hold n_rows rows and nnz entries (float values when f32).
alls_status ensure_csr(alls_csr* out, int64_t n_rows, int64_t nnz, int i64, int f32 = 0);
void free_csr(alls_csr* out);
alls_status hip_fail(hipError_t e, const char* what, std::string* err);
// Pattern cache internals shared with the server (pattern.cpp).
alls_status pattern_build(Workspace& ws, const alls_mesh* mesh, const alls_opts& o, hipStream_t s, int device,
                          alls_pattern** out_pat, alls_csr* out, alls_stats* st, std::string* err);
// E_dev (device, per element) or, when null, the scalar E_scalar.
alls_status pattern_values(Workspace& ws, alls_pattern* p, const double* E_dev, double E_scalar, alls_csr* out,
                           hipStream_t s, alls_stats* st, std::string* err);
void pattern_destroy(alls_pattern* p);
int64_t pattern_bytes(const alls_pattern* p);

}  // namespace alls

// Autotuned launch settings of one (element kind, precision, size class).
struct alls_tuned {
    int32_t block_threads, chunk_slots, grid_per_cu;
    double ms;
};

struct alls_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    alls_stats stats{};
    alls::Workspace ws;
    std::map<uint64_t, alls_tuned> tuned;   // autotuner results by tune_key()
};

namespace alls {
// opts.autotune: copies the tuned block_threads / chunk_slots / grid_per_cu
// of the mesh's kind and size class into *o, tuning on dm first when new.
alls_status apply_tuning(alls_ctx* ctx, const DevMesh& dm, const MeshInfo& mi, alls_opts* o);
// n meshes of one dimension in one packed pipeline on ws/s (alls_assemble_batch);
// tune_ctx (may be null) serves opts.autotune; use_graph replays captured
// graphs (worth it when the same batch shape repeats).
alls_status batch_assemble(Workspace& ws, hipStream_t s, int32_t n, const alls_mesh* meshes, alls_opts o,
                           alls_csr* out, int64_t* row_off, int64_t* nnz_off, alls_stats* stats, std::string* err,
                           alls_ctx* tune_ctx, bool use_graph);
}  // namespace alls
