// assemble.cpp -- context, the A-LLS pipeline driver and CSR ownership.
//
// Replaces the reference's compareAssemblyMethods A-LLS half (Acc:750-838):
// that driver allocates heads/pool per run, publishes the pool through
// __device__ globals and times insert+convert on the legacy stream.  Here the
// pipeline runs on the context's stream with reusable scratch, checks every
// HIP call, detects list-directory overflow and retries with larger chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_internal.h"

namespace alls {

// ------------------------------------------------ caching device allocator
// hipFree synchronises the whole device, which would serialise the server's
// streams.  Freed blocks go to a size-ordered free list and are reused by
// later requests; they are returned to HIP only when an allocation fails or
// the cache exceeds a quarter of device memory (the server's default budget
// leaves that much HBM free).  Every free happens after the
// host has synchronised the work that used the block, so reuse is safe.
namespace {
struct Block {
    size_t bytes;
    int device;
};
std::mutex g_alloc_mu;
std::multimap<size_t, std::pair<void*, int>> g_free;     // size -> (ptr, device)
std::unordered_map<void*, Block> g_live;
size_t g_cached = 0;

size_t round_size(size_t n) {
    if (n < (size_t(1) << 20)) return (n + 4095) & ~size_t(4095);
    return (n + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
}

size_t cache_cap() {
    static size_t cap = 0;
    if (!cap) {
        size_t fr = 0, tot = 0;
        cap = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 4 : (size_t(16) << 30);
    }
    return cap;
}

// Frees cached blocks, largest first, until `need` bytes are returned (all
// of them when need == 0).  Caller holds g_alloc_mu.
void trim_locked(size_t need) {
    size_t freed = 0;
    while (!g_free.empty() && (need == 0 || freed < need)) {
        auto it = std::prev(g_free.end());
        (void)hipFree(it->second.first);
        g_live.erase(it->second.first);
        freed += it->first;
        g_cached -= it->first;
        g_free.erase(it);
    }
}
}  // namespace

void dev_trim() {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    trim_locked(0);
}

// ------------------------------------------------------------ device arenas
// A server reserves its whole HBM budget with one hipMalloc and its worker
// threads carve every buffer (workspaces, outputs, cached patterns) out of
// it: best fit over free ranges, coalesced on free.  The serving loop then
// never calls hipMalloc or hipFree, whose cost grows with the size (hundreds
// of ms for tens of GB) and which block other threads' HIP calls.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::mutex mu;
    std::map<size_t, size_t> free_at;          // offset -> length
    std::multimap<size_t, size_t> free_len;    // length -> offset
    std::unordered_map<size_t, size_t> live;   // offset -> length
};

namespace {
std::mutex g_arena_mu;
std::vector<Arena*> g_arenas;
thread_local Arena* t_arena = nullptr;
// Set for the duration of an alls_assemble_noalloc call: any device
// allocation on this thread then fails instead of happening.
thread_local bool t_no_alloc = false;

void arena_insert_free(Arena* a, size_t off, size_t len) {
    a->free_at[off] = len;
    a->free_len.emplace(len, off);
}
void arena_erase_free(Arena* a, std::map<size_t, size_t>::iterator it) {
    auto r = a->free_len.equal_range(it->second);
    for (auto j = r.first; j != r.second; ++j)
        if (j->second == it->first) {
            a->free_len.erase(j);
            break;
        }
    a->free_at.erase(it);
}

void* arena_alloc(Arena* a, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    std::lock_guard<std::mutex> lk(a->mu);
    auto it = a->free_len.lower_bound(need);
    if (it == a->free_len.end()) return nullptr;
    const size_t off = it->second, len = it->first;
    arena_erase_free(a, a->free_at.find(off));
    if (len > need) arena_insert_free(a, off + need, len - need);
    a->live[off] = need;
    return a->base + off;
}

// True when p belonged to an arena (and is now free in it).
bool arena_free(void* p) {
    Arena* a = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_arena_mu);
        for (Arena* x : g_arenas)
            if (static_cast<char*>(p) >= x->base && static_cast<char*>(p) < x->base + x->size) a = x;
    }
    if (!a) return false;
    std::lock_guard<std::mutex> lk(a->mu);
    size_t off = (size_t)(static_cast<char*>(p) - a->base);
    auto lv = a->live.find(off);
    if (lv == a->live.end()) return true;   // not a live block: ignore
    size_t len = lv->second;
    a->live.erase(lv);
    auto nx = a->free_at.lower_bound(off);
    if (nx != a->free_at.end() && nx->first == off + len) {
        len += nx->second;
        arena_erase_free(a, nx);
    }
    auto pv = a->free_at.lower_bound(off);
    if (pv != a->free_at.begin()) {
        --pv;
        if (pv->first + pv->second == off) {
            off = pv->first;
            len += pv->second;
            arena_erase_free(a, pv);
        }
    }
    arena_insert_free(a, off, len);
    return true;
}
}  // namespace

Arena* arena_create(size_t bytes) {
    bytes &= ~size_t(255);
    if (bytes == 0) return nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    Arena* a = new Arena();
    a->base = static_cast<char*>(p);
    a->size = bytes;
    arena_insert_free(a, 0, bytes);
    std::lock_guard<std::mutex> lk(g_arena_mu);
    g_arenas.push_back(a);
    return a;
}

void arena_destroy(Arena* a) {
    if (!a) return;
    {
        std::lock_guard<std::mutex> lk(g_arena_mu);
        g_arenas.erase(std::remove(g_arenas.begin(), g_arenas.end(), a), g_arenas.end());
    }
    (void)hipFree(a->base);
    delete a;
}

void arena_bind(Arena* a) { t_arena = a; }

void no_alloc_scope(bool on) { t_no_alloc = on; }

hipError_t dev_alloc(void** out, size_t bytes) {
    if (t_no_alloc) {
        *out = nullptr;
        return hipErrorOutOfMemory;
    }
    if (t_arena) {
        // Arena first; a request it cannot fit falls back to the cache below.
        void* p = arena_alloc(t_arena, std::max<size_t>(bytes, 256));
        if (p) {
            *out = p;
            return hipSuccess;
        }
    }
    bytes = round_size(std::max<size_t>(bytes, 256));
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        // Smallest cached block that holds the request and wastes at most half.
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        for (auto it = g_free.lower_bound(bytes); it != g_free.end() && it->first <= 2 * bytes; ++it) {
            if (it->second.second != dev) continue;
            *out = it->second.first;
            g_cached -= it->first;
            g_free.erase(it);
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_alloc_mu);
            trim_locked(2 * bytes);
        }
        e = hipMalloc(out, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            dev_trim();
            e = hipMalloc(out, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            return e;
        }
    }
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_live[*out] = Block{bytes, dev};
    return hipSuccess;
}

void dev_free(void* p) {
    if (!p) return;
    if (arena_free(p)) return;
    std::unique_lock<std::mutex> lk(g_alloc_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        lk.unlock();
        (void)hipFree(p);
        return;
    }
    if (g_cached + it->second.bytes > cache_cap()) trim_locked(it->second.bytes);
    g_cached += it->second.bytes;
    g_free.emplace(it->second.bytes, std::make_pair(p, it->second.device));
}

hipError_t DevBuf::ensure(size_t need) {
    if (need == 0) need = 256;
    if (need <= bytes) return hipSuccess;
    release();
    const size_t nb = need + need / 16 + 256;
    hipError_t e = dev_alloc(&p, nb);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    bytes = nb;
    return hipSuccess;
}

void DevBuf::release() {
    dev_free(p);
    p = nullptr;
    bytes = 0;
}

hipError_t Workspace::init() {
    if (!host_misc) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&host_misc), 64 * sizeof(int64_t), hipHostMallocDefault);
        if (e != hipSuccess) {
            host_misc = nullptr;
            return e;
        }
    }
    if (!events) {
        for (auto& x : ev) {
            hipError_t e = hipEventCreate(&x);
            if (e != hipSuccess) return e;
        }
        events = true;
    }
    return hipSuccess;
}

// Host copy into the pinned staging buffer.  Large inputs are split over a
// few threads: one thread's memcpy (about 5-10 GB/s) would otherwise cost
// more than the GPU work of the assembly it feeds.
static void staged_copy(void* dst, const void* src, size_t n) {
    constexpr size_t kSplit = size_t(32) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t parts = std::min<size_t>(std::min<size_t>(8, hw), n / kSplit);
    if (parts < 2) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t step = ((n + parts - 1) / parts + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    for (size_t k = 1; k < parts && k * step < n; ++k)
        th.emplace_back([=] {
            std::memcpy(static_cast<char*>(dst) + k * step, static_cast<const char*>(src) + k * step,
                        std::min(step, n - k * step));
        });
    std::memcpy(dst, src, std::min(step, n));
    for (auto& t : th) t.join();
}

hipError_t Workspace::staging(size_t total, char** host) {
    *host = nullptr;
    // The previous upload from this buffer has completed: every pipeline
    // synchronises its stream before returning.
    if (total > pinned_bytes) {
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_bytes = 0;
        const size_t nb = total + total / 4;
        hipError_t e = hipHostMalloc(&pinned, nb, hipHostMallocDefault);
        if (e != hipSuccess) {
            pinned = nullptr;
            (void)hipGetLastError();
            return e;
        }
        pinned_bytes = nb;
    }
    *host = static_cast<char*>(pinned);
    return hipSuccess;
}

hipError_t Workspace::upload(void* const* dst, const void* const* src, const size_t* bytes, int n, hipStream_t s) {
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += (bytes[i] + 255) & ~size_t(255);
    if (total == 0) return hipSuccess;
    // The previous upload from this buffer has completed: every pipeline
    // synchronises its stream before returning.
    if (total > pinned_bytes) {
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_bytes = 0;
        const size_t nb = total + total / 4;
        hipError_t e = hipHostMalloc(&pinned, nb, hipHostMallocDefault);
        if (e != hipSuccess) {
            pinned = nullptr;
            (void)hipGetLastError();
            return e;
        }
        pinned_bytes = nb;
    }
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        if (!bytes[i]) continue;
        char* stage = static_cast<char*>(pinned) + off;
        staged_copy(stage, src[i], bytes[i]);
        hipError_t e = hipMemcpyAsync(dst[i], stage, bytes[i], hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        off += (bytes[i] + 255) & ~size_t(255);
    }
    return hipSuccess;
}

size_t Workspace::device_bytes() const {
    size_t b = 0;
    for (const DevBuf* x : {&hdr, &keys, &geid, &vals, &misc, &u, &big, &tiles, &coords, &conn, &types, &E, &bdesc, &bsplit})
        b += x->bytes;
    return b;
}

void Workspace::release_device() {
    graph.reset();
    for (DevBuf* b : {&hdr, &keys, &geid, &vals, &misc, &u, &big, &tiles, &coords, &conn, &types, &E, &bdesc, &bsplit})
        b->release();
}

void Workspace::destroy() {
    graph.reset();
    for (DevBuf* b : {&hdr, &keys, &geid, &vals, &misc, &u, &big, &tiles, &coords, &conn, &types, &E, &bdesc, &bsplit})
        b->release();
    if (host_misc) (void)hipHostFree(host_misc);
    host_misc = nullptr;
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
    if (events)
        for (auto& x : ev) (void)hipEventDestroy(x);
    events = false;
}

alls_status hip_fail(hipError_t e, const char* what, std::string* err) {
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ALLS_E_OOM : ALLS_E_HIP;
}

#define HIP_TRY(call, what)                                   \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail(e_, what, err); \
    } while (0)

alls_status classify_mesh(const alls_mesh* m, const alls_opts& o, MeshInfo* mi, std::string* err) {
    if (!m || !mi) {
        if (err) *err = "null mesh";
        return ALLS_E_INVALID;
    }
    if (m->n_nodes < 0 || m->n_elems < 0 || m->n_nodes >= (int64_t)INT32_MAX / 3 ||
        m->n_elems >= (int64_t)INT32_MAX) {
        if (err) *err = "n_nodes/n_elems out of range";
        return ALLS_E_INVALID;
    }
    if ((m->n_elems > 0 && (!m->conn || !m->coords)) || (m->n_nodes > 0 && !m->coords)) {
        if (err) *err = "missing coords/conn";
        return ALLS_E_INVALID;
    }
    if (m->mem != ALLS_MEM_HOST && m->mem != ALLS_MEM_DEVICE) {
        if (err) *err = "bad mem";
        return ALLS_E_INVALID;
    }
    if (o.val_f32 && m->dim != 3) {
        if (err) *err = "FP32 output values (opts.val_f32) are for Hex20 meshes";
        return ALLS_E_UNSUPPORTED;
    }
    if (m->dim == 3) {
        if (m->nodes_per_elem != 20 || (m->elem_type == nullptr && m->uniform_type != ALLS_HEX20)) {
            if (err) *err = "3D meshes must be Hex20 with nodes_per_elem = 20";
            return ALLS_E_INVALID;
        }
        mi->kind = 3;
        mi->dpn = 3;
        mi->npe_max = 20;
        mi->f32 = o.precision == ALLS_PREC_F32_MFMA;
    } else if (m->dim == 2) {
        if (m->elem_type) {
            if (m->nodes_per_elem != 4) {
                if (err) *err = "mixed 2D meshes need nodes_per_elem = 4 (T3 padded with -1)";
                return ALLS_E_INVALID;
            }
            mi->kind = 2;
            mi->npe_max = 4;
        } else if (m->uniform_type == ALLS_Q4) {
            if (m->nodes_per_elem != 4) {
                if (err) *err = "Q4 meshes need nodes_per_elem = 4";
                return ALLS_E_INVALID;
            }
            mi->kind = 0;
            mi->npe_max = 4;
        } else if (m->uniform_type == ALLS_T3) {
            if (m->nodes_per_elem < 3) {
                if (err) *err = "T3 meshes need nodes_per_elem >= 3";
                return ALLS_E_INVALID;
            }
            mi->kind = 1;
            mi->npe_max = 3;
        } else {
            if (err) *err = "2D meshes must be Q4, T3 or mixed";
            return ALLS_E_INVALID;
        }
        mi->dpn = 2;
        mi->f32 = false;
    } else {
        if (err) *err = "dim must be 2 or 3";
        return ALLS_E_INVALID;
    }
    if (m->nu <= -1.0 || m->nu >= 0.5 || (m->dim == 2 && m->nu >= 1.0)) {
        if (err) *err = "nu out of range";
        return ALLS_E_INVALID;
    }
    mi->entries_bound = m->n_elems * (int64_t)mi->npe_max * mi->npe_max;
    return ALLS_OK;
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

alls_status stage_mesh(Workspace& ws, const alls_mesh* m, const MeshInfo& mi, DevMesh* dm, hipStream_t s,
                       std::string* err) {
    DevMesh d{};
    d.n_nodes = m->n_nodes;
    d.n_elems = m->n_elems;
    d.dim = m->dim;
    d.npe = m->nodes_per_elem;
    d.uniform_type = m->uniform_type;
    d.E = m->E;
    d.nu = m->nu;
    d.thickness = m->thickness;
    const size_t cb = (size_t)m->n_nodes * m->dim * sizeof(double);
    const size_t kb = (size_t)m->n_elems * m->nodes_per_elem * sizeof(int32_t);
    const size_t tb = m->elem_type ? (size_t)m->n_elems : 0;
    const size_t eb = m->E_per_elem ? (size_t)m->n_elems * sizeof(double) : 0;
    if (m->mem == ALLS_MEM_HOST) {
        // Host mesh: one pinned staging pass, four async copies.
        HIP_TRY(ws.coords.ensure(cb), "alloc coords");
        HIP_TRY(ws.conn.ensure(kb), "alloc conn");
        if (tb) HIP_TRY(ws.types.ensure(tb), "alloc types");
        if (eb) HIP_TRY(ws.E.ensure(eb), "alloc E");
        void* dst[4] = {ws.coords.p, ws.conn.p, ws.types.p, ws.E.p};
        const void* src[4] = {m->coords, m->conn, m->elem_type, m->E_per_elem};
        const size_t bytes[4] = {cb, kb, tb, eb};
        HIP_TRY(ws.upload(dst, src, bytes, 4, s), "upload mesh");
        d.coords = cb ? ws.coords.as<double>() : nullptr;
        d.conn = kb ? ws.conn.as<int32_t>() : nullptr;
        d.types = tb ? ws.types.as<uint8_t>() : nullptr;
        d.E_elem = eb ? ws.E.as<double>() : nullptr;
    } else {
        // Device inputs are used in place when aligned for vector loads.
        if (aligned16(m->coords) || !cb) {
            d.coords = m->coords;
        } else {
            HIP_TRY(ws.coords.ensure(cb), "alloc coords");
            HIP_TRY(hipMemcpyAsync(ws.coords.p, m->coords, cb, hipMemcpyDeviceToDevice, s), "copy coords");
            d.coords = ws.coords.as<double>();
        }
        if (aligned16(m->conn) || !kb) {
            d.conn = m->conn;
        } else {
            HIP_TRY(ws.conn.ensure(kb), "alloc conn");
            HIP_TRY(hipMemcpyAsync(ws.conn.p, m->conn, kb, hipMemcpyDeviceToDevice, s), "copy conn");
            d.conn = ws.conn.as<int32_t>();
        }
        d.types = m->elem_type;
        d.E_elem = m->E_per_elem;
    }
    (void)mi;
    *dm = d;
    return ALLS_OK;
}

alls_status ensure_csr(alls_csr* out, int64_t n_rows, int64_t nnz, int i64, int f32) {
    if (out->cap_rows < n_rows + 1 || !out->row_ptr) {
        dev_free(out->row_ptr);
        out->row_ptr = nullptr;
        out->cap_rows = 0;
        if (dev_alloc(&out->row_ptr, (size_t)(n_rows + 1) * sizeof(int64_t)) != hipSuccess) {
            out->row_ptr = nullptr;
            return ALLS_E_OOM;
        }
        out->cap_rows = n_rows + 1;
        out->pattern_id = 0;
    }
    if (out->cap_nnz < nnz || !out->col_ind || !out->val || out->val_is_f32 != f32) {
        dev_free(out->col_ind);
        dev_free(out->val);
        out->col_ind = nullptr;
        out->val = nullptr;
        out->cap_nnz = 0;
        const int64_t n = std::max<int64_t>(nnz, 1);
        if (dev_alloc(reinterpret_cast<void**>(&out->col_ind), (size_t)n * sizeof(int32_t)) != hipSuccess ||
            dev_alloc(reinterpret_cast<void**>(&out->val), (size_t)n * (f32 ? sizeof(float) : sizeof(double))) !=
                hipSuccess) {
            return ALLS_E_OOM;
        }
        out->cap_nnz = n;
        out->pattern_id = 0;
    }
    out->n_rows = n_rows;
    out->nnz = nnz;
    out->row_ptr_is_i64 = i64;
    out->val_is_f32 = f32;
    return ALLS_OK;
}

void free_csr(alls_csr* out) {
    if (!out) return;
    dev_free(out->row_ptr);
    dev_free(out->col_ind);
    dev_free(out->val);
    std::memset(out, 0, sizeof(*out));
}

int default_chunk(int kind) {
    switch (kind) {
        case 0: return 16;   // Q4: a structured interior node row gets 4 elements x 4 entries
        case 1: return 24;   // T3: up to 8 triangles x 3 entries around a grid node
        case 2: return 24;   // mixed Q4/T3
        // Hex20: 20 entries per element; vertex rows of a structured brick
        // take 160, edge rows 80.  One chunk per vertex row means no
        // overflow chunk to take and publish in the insertion and no
        // directory lookup in the merge (80^3 FP32 insertion 4.25 -> 2.98 ms).
        default: return 160;
    }
}

int64_t pool_chunks(int64_t n_nodes, int64_t entries_bound, int C, int kind, bool full) {
    const int64_t bound = (entries_bound + C - 1) / C + 1;
    if (full) return n_nodes + bound;
    // Hex20 chunks of 160 hold every row of a conforming brick: the reserve
    // is for rows of nodes in more than 8 elements.
    const int64_t ovf = kind == 3 && C >= 160 ? n_nodes / 8 + 1024 : bound / 4 + n_nodes / 4 + 1024;
    return n_nodes + std::min(bound, ovf);
}
static int small_group(int kind) { return kind == 0 ? 16 : kind == 3 ? 0 : 32; }

static float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

// Grid cap of the insert kernel: opts.grid_per_cu blocks per CU (persistent,
// grid-stride), or 0 for one thread (2D) / one block (Hex20) per element.
static int64_t insert_max_blocks(int grid_per_cu) {
    if (grid_per_cu <= 0) return 0;
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus = n;
    }
    return (int64_t)grid_per_cu * cus;
}

// State carried across the pipeline's one host synchronisation.
struct Pipeline {
    int C = 0, G = 0;
    int64_t ntiles = 0, cap_chunks = 0;
    bool full_pool = false;
    int retries = 0;
    ListPool lp{};
    int64_t total = 0;     // unique (row node, column node) pairs
    int32_t max_big = 0;   // longest row of the wave-per-row kernels
    int64_t list_bytes = 0;   // headers, keys, group ids and values of this call's pool
    int ng = 0;               // list layout (ListPool::ng), fixed for the call
};

static bool t3_line_groups(int64_t n_nodes);
static bool hex20_segments(int C);

// Chunk size, small-row lanes, scan tiles and list layout of a call.
static void init_pipeline(Pipeline& pl, const alls_opts& o, const MeshInfo& mi, int64_t n_nodes) {
    pl = Pipeline{};
    pl.C = o.chunk_slots > 0 ? o.chunk_slots : default_chunk(mi.kind);
    pl.C = std::min(std::max(pl.C, 4), ALLS_MAX_CHUNK);
    if (mi.kind == 0) pl.C = std::max(4, pl.C / 4 * 4);   // whole slot groups per chunk
    if (mi.kind == 1) pl.C = std::max(3, pl.C / 3 * 3);   // whole T3 line groups per chunk
    pl.G = small_group(mi.kind);
    pl.ntiles = scan_tile_count(n_nodes);
    pl.ng = mi.kind == 0 ? 4 : (mi.kind == 1 && t3_line_groups(n_nodes)) ? 3
          : (mi.kind == 3 && hex20_segments(pl.C)) ? SEG20 : 0;
}

// Hex20 lists carry one element id per 20-slot strip when chunks hold whole
// strips; other chunk sizes key every entry.  ALLS_HEX20_KEYED=1 in the
// environment forces per-entry keys (a diagnostic switch: the tests run both
// layouts).
static bool hex20_segments(int C) {
    if (C % SEG20 != 0) return false;
    const char* f = std::getenv("ALLS_HEX20_KEYED");
    return !(f && f[0] && std::strcmp(f, "0") != 0);
}

// T3 line groups sort 32-bit keys (column << 5 | element-order slot), so
// they take meshes whose node ids fit 27 bits; larger triangle meshes key
// every entry.  ALLS_T3_KEYED=1 in the environment forces the keyed layout
// (a diagnostic switch: the tests run both layouts).
static bool t3_line_groups(int64_t n_nodes) {
    if (n_nodes >= (int64_t(1) << 27)) return false;
    const char* f = std::getenv("ALLS_T3_KEYED");
    return !(f && f[0] && std::strcmp(f, "0") != 0);
}

// List pool for pl.C (pool_chunks).  The exact bound of overflow chunks
// (entries / C) is rarely approached, since most rows fit their home chunk:
// 2D lists reserve a quarter of it plus a chunk per 4 rows, Hex20 lists of
// 160-entry chunks a chunk per 8 rows, and the full bound is taken after a
// pool overflow.
static alls_status alloc_lists(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, Pipeline& pl,
                               std::string* err) {
    const int B = mi.dpn * mi.dpn;
    const size_t vsz = mi.f32 ? sizeof(float) : sizeof(double);
    const int64_t n_nodes = dm.n_nodes;
    pl.cap_chunks = pool_chunks(n_nodes, mi.entries_bound, pl.C, mi.kind, pl.full_pool);
    // Q4 lists carry one element id per group of 4 slots (geid); the
    // per-entry keys are written only for the rows the wave-per-row kernels
    // take.  T3 lists are line groups: each group of 3 slots in one 128-byte
    // line of vals with its key block, and no key array.  Hex20 lists carry
    // one element id per 20-slot strip (geid) and no key array.  Mixed lists
    // (and Hex20 lists with chunks that are not whole strips) key every
    // entry.  (T3 with element ids per group paid a
    // scattered connectivity load per group in the fused pass, 1.34 ms
    // against 1.04 ms keyed at 1500x1500 with permuted node ids; the line
    // groups carry the columns themselves.)
    const int ng = pl.ng;
    const size_t keys_bytes = ng == 3 || ng == SEG20 ? 0 : (size_t)pl.cap_chunks * pl.C * sizeof(int2);
    const size_t geid_bytes = ng == 4 || ng == SEG20 ? (size_t)pl.cap_chunks * (pl.C / ng) * sizeof(int32_t) : 0;
    const size_t vals_bytes = ng == 3 ? (size_t)pl.cap_chunks * (pl.C / 3) * 128
                                      : (size_t)pl.cap_chunks * pl.C * B * vsz;
    pl.list_bytes = (int64_t)((size_t)std::max<int64_t>(n_nodes, 1) * ALLS_DIR * sizeof(int32_t) + keys_bytes +
                              geid_bytes + vals_bytes);
    if (ws.ext) {
        // Fixed scratch (alls_assemble_noalloc): every array is carved from
        // it, 256-byte aligned, and nothing is allocated.
        const size_t nn1 = (size_t)std::max<int64_t>(n_nodes, 1);
        const size_t need[] = {nn1 * ALLS_DIR * sizeof(int32_t), keys_bytes, geid_bytes, vals_bytes,
                               MISC_SIZE * sizeof(int32_t), nn1 * sizeof(int32_t), nn1 * sizeof(int32_t),
                               (size_t)(pl.ntiles + 1) * sizeof(int64_t)};
        DevBuf* bufs[] = {&ws.hdr, &ws.keys, &ws.geid, &ws.vals, &ws.misc, &ws.u, &ws.big, &ws.tiles};
        size_t off = 0;
        for (size_t k = 0; k < sizeof(need) / sizeof(need[0]); ++k) off += (need[k] + 255) & ~size_t(255);
        if (off > ws.ext_bytes) {
            if (err)
                *err = "the mesh needs " + std::to_string(off) + " bytes of list and consolidation scratch; "
                       "the fixed scratch holds " + std::to_string(ws.ext_bytes);
            return ALLS_E_UNSUPPORTED;
        }
        off = 0;
        for (size_t k = 0; k < sizeof(need) / sizeof(need[0]); ++k) {
            bufs[k]->p = need[k] ? ws.ext + off : nullptr;
            bufs[k]->bytes = need[k];
            off += (need[k] + 255) & ~size_t(255);
        }
    }
    HIP_TRY(ws.hdr.ensure((size_t)std::max<int64_t>(n_nodes, 1) * ALLS_DIR * sizeof(int32_t)), "alloc list headers");
    if (keys_bytes) HIP_TRY(ws.keys.ensure(keys_bytes), "alloc list keys");
    if (geid_bytes) HIP_TRY(ws.geid.ensure(geid_bytes), "alloc list groups");
    HIP_TRY(ws.vals.ensure(vals_bytes), "alloc list values");
    HIP_TRY(ws.misc.ensure(MISC_SIZE * sizeof(int32_t)), "alloc misc");
    HIP_TRY(ws.u.ensure((size_t)std::max<int64_t>(n_nodes, 1) * sizeof(int32_t)), "alloc row counts");
    HIP_TRY(ws.big.ensure((size_t)std::max<int64_t>(n_nodes, 1) * sizeof(int32_t)), "alloc big-row list");
    HIP_TRY(ws.tiles.ensure((size_t)(pl.ntiles + 1) * sizeof(int64_t)), "alloc scan tiles");
    pl.lp.hdr = ws.hdr.as<int32_t>();
    pl.lp.keys = keys_bytes ? ws.keys.as<int2>() : nullptr;
    pl.lp.vals = ws.vals.p;
    pl.lp.misc = ws.misc.as<int32_t>();
    pl.lp.n_nodes = n_nodes;
    pl.lp.cap_chunks = pl.cap_chunks;
    pl.lp.C = pl.C;
    pl.lp.B = B;
    pl.lp.cmagic = chunk_magic(pl.C);
    pl.lp.long_row = pl.G > 0 ? pl.G : INT32_MAX;
    pl.lp.ng = ng;
    pl.lp.gpc = ng ? pl.C / ng : 0;
    pl.lp.geid = geid_bytes ? ws.geid.as<int32_t>() : nullptr;
    pl.lp.conn = dm.conn;
    pl.lp.n_elems = dm.n_elems;
    pl.lp.npe = dm.npe;
    return ALLS_OK;
}

// Header reset, integration + insertion, count, scan totals, and the
// read-back of the flags and the total into pinned memory.  `events`
// records the stage events (not under graph capture).
static alls_status enqueue_insert(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o,
                                  const Pipeline& pl, hipStream_t s, bool events, std::string* err) {
    const int64_t maxb = insert_max_blocks(o.grid_per_cu);
    if (events) HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
    launch_fill_i32(pl.lp.hdr, dm.n_nodes * ALLS_DIR, 0, s);
    launch_fill_i32(pl.lp.misc, MISC_SIZE, 0, s);
    if (events) HIP_TRY(hipEventRecord(ws.ev[1], s), "event");
    if (mi.kind == 3) launch_insert_hex20(dm, pl.lp, mi.f32 ? 1 : 0, maxb, s);
    else launch_insert_2d(dm, pl.lp, mi.kind, o.block_threads, maxb, s);
    HIP_TRY(hipGetLastError(), "launch insert");
    if (events) HIP_TRY(hipEventRecord(ws.ev[2], s), "event");
    return ALLS_OK;
}

// Count pass, scan totals and the read-back of the flags and the total.
static alls_status enqueue_readback(Workspace& ws, const Pipeline& pl, hipStream_t s, std::string* err) {
    HIP_TRY(hipMemcpyAsync(ws.host_misc, pl.lp.misc, MISC_SIZE * sizeof(int32_t), hipMemcpyDeviceToHost, s),
            "read flags");
    HIP_TRY(hipMemcpyAsync(ws.host_misc + 32, ws.tiles.as<int64_t>() + pl.ntiles, sizeof(int64_t),
                           hipMemcpyDeviceToHost, s), "read total");
    return ALLS_OK;
}

static alls_status enqueue_count(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const Pipeline& pl,
                                 hipStream_t s, bool events, std::string* err, bool readback = true) {
    launch_count(pl.lp, pl.G, mi.f32 ? 1 : 0, ws.u.as<int32_t>(), ws.big.as<int32_t>(), s);
    HIP_TRY(hipGetLastError(), "launch count");
    if (events) HIP_TRY(hipEventRecord(ws.ev[3], s), "event");
    launch_scan_total(ws.u.as<int32_t>(), dm.n_nodes, ws.tiles.as<int64_t>(), pl.ntiles, s);
    HIP_TRY(hipGetLastError(), "launch scan");
    if (events) HIP_TRY(hipEventRecord(ws.ev[4], s), "event");
    return readback ? enqueue_readback(ws, pl, s, err) : ALLS_OK;
}

// Header reset, integration + insertion, count, scan totals, and the
// read-back of the flags and the total into pinned memory.  `events`
// records the stage events (not under graph capture).
static alls_status enqueue_front(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o,
                                 const Pipeline& pl, hipStream_t s, bool events, std::string* err) {
    alls_status rs = enqueue_insert(ws, dm, mi, o, pl, s, events, err);
    if (rs != ALLS_OK) return rs;
    return enqueue_count(ws, dm, mi, pl, s, events, err);
}

// Reads the flags after the stream has synchronised.  *retry: run the front
// half again with pl updated (larger chunks or the full pool).
static alls_status check_front(Workspace& ws, Pipeline& pl, bool* retry, std::string* err) {
    *retry = false;
    const int32_t* f = reinterpret_cast<const int32_t*>(ws.host_misc);
    if (f[MISC_ERR + ERR_INVALID_MESH]) {
        if (err) *err = "mesh has an element with a node id out of range or a wrong element type";
        return ALLS_E_INVALID;
    }
    if (f[MISC_ERR + ERR_SPIN_TIMEOUT]) {
        if (err) *err = "list chunk publication timed out";
        return ALLS_E_HIP;
    }
    const int32_t need = f[MISC_ERR + ERR_DIR_OVERFLOW];
    if (need > 0 || f[MISC_ERR + ERR_POOL_EXHAUSTED] || f[MISC_ERR + ERR_ROW_TOO_BIG]) {
        if (++pl.retries > 6) {
            if (err) *err = "list directory overflow after retries";
            return ALLS_E_POOL_OVERFLOW;
        }
        if (need > 0 || f[MISC_ERR + ERR_ROW_TOO_BIG]) {
            int64_t nc = 2 * (int64_t)pl.C;
            while (nc * ALLS_DIR < need) nc *= 2;
            if (nc > ALLS_MAX_CHUNK) {
                if (err)
                    *err = "a node row needs " + std::to_string(need) + " list entries; at most " +
                           std::to_string(long_row_limit()) + " are supported";
                return ALLS_E_UNSUPPORTED;
            }
            pl.C = (int)nc;
        } else {
            pl.full_pool = true;
        }
        *retry = true;
        return ALLS_OK;
    }
    pl.max_big = f[MISC_MAX_BIG];
    if (pl.max_big > long_row_limit()) {
        if (err)
            *err = "a node row has " + std::to_string(pl.max_big) + " list entries; at most " +
                   std::to_string(long_row_limit()) + " are supported";
        return ALLS_E_UNSUPPORTED;
    }
    pl.total = ws.host_misc[32];
    return ALLS_OK;
}

// row_ptr from the scan, then the per-row merge and write.
static alls_status enqueue_back(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const Pipeline& pl,
                                alls_csr* out, hipStream_t s, bool events, std::string* err) {
    if (events) HIP_TRY(hipEventRecord(ws.ev[5], s), "event");
    launch_scan_apply(ws.u.as<int32_t>(), dm.n_nodes, ws.tiles.as<int64_t>(), mi.dpn, out->row_ptr,
                      out->row_ptr_is_i64, s);
    HIP_TRY(hipGetLastError(), "launch scan apply");
    if (events) HIP_TRY(hipEventRecord(ws.ev[7], s), "event");
    launch_write(pl.lp, pl.G, mi.f32 ? 1 : 0, mi.dpn, pl.max_big, ws.big.as<int32_t>(), out->row_ptr,
                 out->row_ptr_is_i64, out->col_ind, out->val, out->val_is_f32, s);
    HIP_TRY(hipGetLastError(), "launch write");
    if (events) HIP_TRY(hipEventRecord(ws.ev[6], s), "event");
    return ALLS_OK;
}

// The fused small-row pass: insertion, then count + scan + write in one
// launch into an output sized by the entries bound (each entry is at most
// one new column).  *done: the CSR is complete.  *fallback: the lists are
// fine but a row is longer than the register kernel holds; run the separate
// passes on them.  Neither: the lists must be rebuilt (pl updated).
static alls_status run_fused(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o, Pipeline& pl,
                             alls_csr* out, hipStream_t s, std::string* err, bool* done, bool* fallback) {
    *done = *fallback = false;
    const int dpn = mi.dpn;
    const int64_t bound = std::max<int64_t>(mi.entries_bound, 1) * dpn * dpn;
    const int i64 = (o.row_ptr_i64 || bound > (int64_t)INT32_MAX) ? 1 : 0;
    alls_status rs = ensure_csr(out, dm.n_nodes * dpn, bound, i64);
    if (rs != ALLS_OK) {
        if (err) *err = "cannot allocate output CSR";
        return rs;
    }
    const int64_t nt = fused_tile_count(dm.n_nodes, pl.G);
    HIP_TRY(ws.tiles.ensure((size_t)std::max<int64_t>(nt, pl.ntiles + 1) * sizeof(int64_t)), "alloc tile states");
    const bool events = o.timing != 0;
    if (!events) HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
    rs = enqueue_insert(ws, dm, mi, o, pl, s, events, err);
    if (rs != ALLS_OK) return rs;
    launch_fill_i32(ws.tiles.as<int32_t>(), nt * 2, 0, s);
    if (events) {
        HIP_TRY(hipEventRecord(ws.ev[3], s), "event");
        HIP_TRY(hipEventRecord(ws.ev[4], s), "event");
        HIP_TRY(hipEventRecord(ws.ev[5], s), "event");
        HIP_TRY(hipEventRecord(ws.ev[7], s), "event");
    }
    launch_rows_fused(pl.lp, pl.G, ws.tiles.as<int64_t>(), out->row_ptr, i64, out->col_ind, out->val, s);
    HIP_TRY(hipGetLastError(), "launch fused pass");
    HIP_TRY(hipEventRecord(ws.ev[6], s), "event");
    HIP_TRY(hipMemcpyAsync(ws.host_misc, pl.lp.misc, MISC_SIZE * sizeof(int32_t), hipMemcpyDeviceToHost, s),
            "read flags");
    HIP_TRY(hipStreamSynchronize(s), "sync after fused pass");
    int32_t* f = reinterpret_cast<int32_t*>(ws.host_misc);
    const bool row_long = f[MISC_MAX_ROW] > pl.G || f[MISC_ERR + ERR_ROW_TOO_BIG] || f[MISC_ERR + ERR_LOOKBACK];
    if (f[MISC_ERR + ERR_INVALID_MESH] || f[MISC_ERR + ERR_SPIN_TIMEOUT] || f[MISC_ERR + ERR_DIR_OVERFLOW] ||
        f[MISC_ERR + ERR_POOL_EXHAUSTED]) {
        f[MISC_ERR + ERR_ROW_TOO_BIG] = 0;   // a fused-pass flag; the lists are rebuilt anyway
        bool retry = false;
        rs = check_front(ws, pl, &retry, err);
        if (rs != ALLS_OK) return rs;
        if (retry) return ALLS_OK;
    }
    if (row_long) {
        *fallback = true;
        return ALLS_OK;
    }
    pl.total = (int64_t)f[MISC_TOTAL_LO] + ((int64_t)f[MISC_TOTAL_HI] << 31);
    pl.max_big = 0;
    *done = true;
    return ALLS_OK;
}

namespace {
uint64_t mix(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)); }
uint64_t mixp(uint64_t h, const void* p) { return mix(h, (uint64_t)(uintptr_t)p); }
uint64_t mixd(uint64_t h, double d) {
    uint64_t b;
    std::memcpy(&b, &d, sizeof b);
    return mix(h, b);
}

// Every kernel argument of the front half: a captured graph replays them.
uint64_t front_key(const Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o,
                   const Pipeline& pl) {
    uint64_t h = 0x5eed;
    for (const void* p : {(const void*)dm.coords, (const void*)dm.conn, (const void*)dm.types,
                          (const void*)dm.E_elem, (const void*)pl.lp.hdr, (const void*)pl.lp.keys,
                          (const void*)pl.lp.geid, (const void*)pl.lp.vals,
                          (const void*)pl.lp.misc, (const void*)ws.u.p, (const void*)ws.big.p, (const void*)ws.tiles.p,
                          (const void*)ws.host_misc})
        h = mixp(h, p);
    for (int64_t v : {dm.n_nodes, dm.n_elems, (int64_t)dm.npe, (int64_t)dm.dim, (int64_t)dm.uniform_type,
                      (int64_t)mi.kind, (int64_t)mi.f32, (int64_t)pl.C, pl.cap_chunks, pl.ntiles, (int64_t)pl.G,
                      (int64_t)pl.ng, (int64_t)o.block_threads, (int64_t)o.grid_per_cu})
        h = mix(h, (uint64_t)v);
    return mixd(mixd(mixd(h, dm.E), dm.nu), dm.thickness);
}

uint64_t back_key(const Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const Pipeline& pl,
                  const alls_csr* out) {
    uint64_t h = 0xbac;
    for (const void* p : {(const void*)pl.lp.hdr, (const void*)pl.lp.keys, (const void*)pl.lp.geid,
                          (const void*)pl.lp.conn, (const void*)pl.lp.vals, (const void*)pl.lp.misc,
                          (const void*)ws.u.p, (const void*)ws.big.p, (const void*)ws.tiles.p, (const void*)out->row_ptr,
                          (const void*)out->col_ind, (const void*)out->val})
        h = mixp(h, p);
    for (int64_t v : {dm.n_nodes, (int64_t)mi.dpn, (int64_t)mi.f32, (int64_t)pl.C, pl.cap_chunks, pl.ntiles,
                      (int64_t)pl.G, (int64_t)pl.max_big, (int64_t)out->row_ptr_is_i64, (int64_t)dm.npe,
                      (int64_t)mi.kind, (int64_t)pl.ng, dm.n_elems})
        h = mix(h, (uint64_t)v);
    return h;
}
}  // namespace

void GraphSlot::reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key = 0;
}

// Launches the graph of `enqueue` on s, capturing it first when the key
// (every argument the captured kernels hold) changed.
template <class F>
static alls_status graph_launch(GraphSlot& g, uint64_t key, hipStream_t s, F&& enqueue, std::string* err) {
    if (!g.exec || g.key != key) {
        g.reset();
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
        const alls_status st = enqueue();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (st != ALLS_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        HIP_TRY(e, "end graph capture");
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g.exec = nullptr;
            return hip_fail(ei, "instantiate graph", err);
        }
        g.key = key;
    }
    HIP_TRY(hipGraphLaunch(g.exec, s), "graph launch");
    return ALLS_OK;
}

// The whole pipeline with ONE host synchronisation, at the end: the output
// is sized by the entries bound (each list entry is at most one new column),
// so no read-back of the total splits the pipeline, and the write pass's
// long-row kernels are launched for rows of up to `cap` entries: the list
// directory's bound in 2D (at most 256), three 60-entry segment slots for
// Hex20 (a conforming brick's rows hold 160).  After it, one read-back of the
// flags and the total: a list overflow rebuilds the lists and runs again; a
// row longer than those kernels take (flagged, or the count pass's longest
// row above `cap`) is written again with the real bound.  capture: the
// pipeline is one hipGraph (batch path), captured on first use and replayed
// while its arguments are unchanged; otherwise it is launched as is, with the
// stage events when opts.timing is set.  *oom: the bound-sized output could
// not be allocated (the caller falls back to an output of the exact size).
// Plain 2D and Hex20 assemblies run this way: the second synchronisation and
// read-back of the two-phase pipeline cost about 40 us of the 1.41 ms C2 step.
static alls_status run_single(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o, Pipeline& pl,
                              alls_csr* out, hipStream_t s, alls_stats* st, std::string* err, bool capture,
                              bool* oom) {
    const int dpn = mi.dpn;
    const int64_t n_rows = dm.n_nodes * dpn;
    const int64_t bound = std::max<int64_t>(mi.entries_bound, 1) * dpn * dpn;
    const int i64 = (o.row_ptr_i64 || bound > (int64_t)INT32_MAX) ? 1 : 0;
    const bool events = !capture && o.timing;
    int launches = 0;
    int32_t cap = 0;
    if (oom) *oom = false;
    for (;;) {
        alls_status rs = alloc_lists(ws, dm, mi, pl, err);
        if (rs != ALLS_OK) return rs;
        rs = ensure_csr(out, n_rows, bound, i64, o.val_f32 ? 1 : 0);
        if (rs != ALLS_OK) {
            if (oom && rs == ALLS_E_OOM && !capture) {
                *oom = true;
                return ALLS_OK;
            }
            if (err) *err = "cannot allocate output CSR";
            return rs;
        }
        cap = mi.kind == 3 ? 180 : std::min(ALLS_DIR * pl.C, 256);
        pl.max_big = cap;
        if (!events) HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
        if (capture) {
            rs = graph_launch(ws.graph, mix(front_key(ws, dm, mi, o, pl), back_key(ws, dm, mi, pl, out)), s,
                              [&] {
                                  alls_status r = enqueue_insert(ws, dm, mi, o, pl, s, false, err);
                                  if (r == ALLS_OK) r = enqueue_count(ws, dm, mi, pl, s, false, err, false);
                                  if (r == ALLS_OK) r = enqueue_back(ws, dm, mi, pl, out, s, false, err);
                                  if (r == ALLS_OK) r = enqueue_readback(ws, pl, s, err);
                                  return r;
                              },
                              err);
            if (rs != ALLS_OK) return rs;
            ++launches;
        } else {
            rs = enqueue_insert(ws, dm, mi, o, pl, s, events, err);
            if (rs == ALLS_OK) rs = enqueue_count(ws, dm, mi, pl, s, events, err, false);
            if (rs == ALLS_OK) rs = enqueue_back(ws, dm, mi, pl, out, s, events, err);
            if (rs == ALLS_OK) rs = enqueue_readback(ws, pl, s, err);
            if (rs != ALLS_OK) return rs;
        }
        if (!events) HIP_TRY(hipEventRecord(ws.ev[6], s), "event");
        HIP_TRY(hipStreamSynchronize(s), capture ? "sync after graph" : "sync after pipeline");
        int32_t* f = reinterpret_cast<int32_t*>(ws.host_misc);
        const int32_t too_big = f[MISC_ERR + ERR_ROW_TOO_BIG];
        f[MISC_ERR + ERR_ROW_TOO_BIG] = 0;   // a write-pass flag here, not a list overflow
        bool retry = false;
        rs = check_front(ws, pl, &retry, err);
        if (rs != ALLS_OK) return rs;
        if (retry) continue;
        if (too_big || pl.max_big > cap) {
            // Rows the write kernels launched could not take: write again.
            rs = enqueue_back(ws, dm, mi, pl, out, s, events, err);
            if (rs != ALLS_OK) return rs;
            if (!events) HIP_TRY(hipEventRecord(ws.ev[6], s), "event");
            HIP_TRY(hipStreamSynchronize(s), "sync after write");
        }
        break;
    }
    out->n_rows = n_rows;
    out->nnz = pl.total * dpn * dpn;
    out->pattern_id = 0;
    if (st) {
        *st = alls_stats{};
        st->ms_total = ev_ms(ws.ev[0], ws.ev[6]);
        if (events) {
            st->ms_init = ev_ms(ws.ev[0], ws.ev[1]);
            st->ms_insert = ev_ms(ws.ev[1], ws.ev[2]);
            st->ms_count = ev_ms(ws.ev[2], ws.ev[3]);
            st->ms_scan = ev_ms(ws.ev[3], ws.ev[4]) + ev_ms(ws.ev[5], ws.ev[7]);
            st->ms_write = ev_ms(ws.ev[7], ws.ev[6]);
        } else if (capture) {
            st->ms_write = st->ms_total;
        }
        st->n_elems = dm.n_elems;
        st->n_rows = n_rows;
        st->nnz = out->nnz;
        st->list_entries = mi.entries_bound;
        st->chunk_slots = pl.C;
        st->retries = pl.retries;
        st->scratch_bytes = pl.list_bytes;
        const int32_t* f = reinterpret_cast<const int32_t*>(ws.host_misc);
        st->big_rows = f[MISC_BIG_COUNT];
        st->pool_chunks_used = dm.n_nodes + f[MISC_OVF_TOP];
        st->pool_chunks_cap = pl.cap_chunks;
        st->graph_launches = launches;
    }
    return ALLS_OK;
}

alls_status run_alls(Workspace& ws, const DevMesh& dm, const MeshInfo& mi, const alls_opts& o, alls_csr* out,
                     hipStream_t s, alls_stats* st, std::string* err, bool use_graph) {
    HIP_TRY(ws.init(), "workspace init");
    const int dpn = mi.dpn;
    const int64_t n_nodes = dm.n_nodes;
    const int64_t n_rows = n_nodes * dpn;
    Pipeline pl;
    init_pipeline(pl, o, mi, n_nodes);
    if (use_graph) return run_single(ws, dm, mi, o, pl, out, s, st, err, true, nullptr);
    const bool stage_events = o.timing != 0;
    // T3 line groups whose rows all fit the register kernel take the fused
    // pass (count + scan + write in one launch), into an output sized by the
    // entries bound; if a row turns out longer, the separate passes run on
    // the same lists.  A count pass over line groups touches every line of a
    // row for its columns (C3 T3 1500x1500: count 0.40 + write 0.65 ms against
    // 1.00 ms fused).  Keyed lists (mixed meshes) count from their compact
    // keys and take the separate passes (C3 mixed: count 0.23 + write 0.62
    // against 1.00 ms fused); so does Q4, whose quad count takes 0.09 ms.
    bool try_fused = !ws.ext && pl.ng == 3 && pl.G == 32 && n_nodes > 0;
    // row_ptr is int32 unless nnz reaches 2^31 (or int64 is asked for), so an
    // entries bound past 2^31 takes the two-phase pipeline, sized exactly.
    const bool bound_fits = o.row_ptr_i64 || std::max<int64_t>(mi.entries_bound, 1) * dpn * dpn <= (int64_t)INT32_MAX;
    if (!ws.ext && !try_fused && bound_fits) {
        bool oom = false;
        const alls_status rs1 = run_single(ws, dm, mi, o, pl, out, s, st, err, false, &oom);
        if (rs1 != ALLS_OK || !oom) return rs1;
        // No room for the bound-sized output: the two-phase pipeline below
        // sizes it exactly.
        init_pipeline(pl, o, mi, n_nodes);
    }
    bool fused_done = false;
    if (!stage_events) HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
    for (;;) {
        alls_status rs = alloc_lists(ws, dm, mi, pl, err);
        if (rs != ALLS_OK) return rs;
        if (try_fused) {
            bool fallback = false;
            rs = run_fused(ws, dm, mi, o, pl, out, s, err, &fused_done, &fallback);
            if (rs != ALLS_OK) return rs;
            if (fused_done) break;
            if (!fallback) continue;   // list retry (larger chunks or the full pool)
            try_fused = false;
            // Separate passes on the lists just built.
            launch_fill_i32(pl.lp.misc, MISC_SIZE, 0, s);
            rs = enqueue_count(ws, dm, mi, pl, s, stage_events, err);
            if (rs != ALLS_OK) return rs;
        } else {
            rs = enqueue_front(ws, dm, mi, o, pl, s, stage_events, err);
            if (rs != ALLS_OK) return rs;
        }
        HIP_TRY(hipStreamSynchronize(s), "sync after scan");
        bool retry = false;
        rs = check_front(ws, pl, &retry, err);
        if (rs != ALLS_OK) return rs;
        if (!retry) break;
    }
    const int64_t nnz = pl.total * dpn * dpn;
    alls_status rs = ALLS_OK;
    if (fused_done) {
        out->nnz = nnz;
    } else if (ws.ext) {
        // The caller's output buffers: checked, never reallocated.
        out->n_rows = n_rows;
        out->nnz = nnz;
        if (out->val_is_f32 != (o.val_f32 ? 1 : 0)) {
            if (err) *err = "opts->val_f32 and out->val_is_f32 disagree";
            return ALLS_E_INVALID;
        }
        if (out->cap_rows < n_rows + 1 || out->cap_nnz < nnz || (!out->row_ptr_is_i64 && nnz > (int64_t)INT32_MAX)) {
            if (err)
                *err = "output buffers too small: need " + std::to_string(n_rows + 1) + " row_ptr entries and " +
                       std::to_string(nnz) + " nonzeros (nnz returned in the CSR)";
            return ALLS_E_INVALID;
        }
    } else {
        const int i64 = (o.row_ptr_i64 || nnz > (int64_t)INT32_MAX) ? 1 : 0;
        rs = ensure_csr(out, n_rows, nnz, i64, o.val_f32 ? 1 : 0);
        if (rs != ALLS_OK) {
            if (err) *err = "cannot allocate output CSR";
            return rs;
        }
    }
    if (fused_done) {
        // nothing left to launch
    } else {
        rs = enqueue_back(ws, dm, mi, pl, out, s, stage_events, err);
        if (rs != ALLS_OK) return rs;
        if (!stage_events) HIP_TRY(hipEventRecord(ws.ev[6], s), "event");
    }
    out->pattern_id = 0;
    HIP_TRY(hipStreamSynchronize(s), "sync after write");
    if (st) {
        *st = alls_stats{};
        if (stage_events) {
            st->ms_init = ev_ms(ws.ev[0], ws.ev[1]);
            st->ms_insert = ev_ms(ws.ev[1], ws.ev[2]);
            st->ms_count = ev_ms(ws.ev[2], ws.ev[3]);
            // The scan's totals, then its application to row_ptr (ahead of the write).
            st->ms_scan = ev_ms(ws.ev[3], ws.ev[4]) + ev_ms(ws.ev[5], ws.ev[7]);
            st->ms_write = ev_ms(ws.ev[7], ws.ev[6]);   // merge/write kernels alone
        }
        st->ms_total = ev_ms(ws.ev[0], ws.ev[6]);   // includes the flag read-back between the halves
        st->n_elems = dm.n_elems;
        st->n_rows = n_rows;
        st->nnz = nnz;
        st->list_entries = mi.entries_bound;
        st->chunk_slots = pl.C;
        st->retries = pl.retries;
        st->scratch_bytes = pl.list_bytes;
        const int32_t* f = reinterpret_cast<const int32_t*>(ws.host_misc);
        st->big_rows = f[MISC_BIG_COUNT];
        st->pool_chunks_used = n_nodes + f[MISC_OVF_TOP];
        st->pool_chunks_cap = pl.cap_chunks;
    }
    return ALLS_OK;
}

}  // namespace alls

using namespace alls;

extern "C" {

const char* alls_status_str(alls_status s) {
    switch (s) {
        case ALLS_OK: return "ok";
        case ALLS_E_INVALID: return "invalid argument or mesh";
        case ALLS_E_OOM: return "out of memory";
        case ALLS_E_POOL_OVERFLOW: return "list pool overflow";
        case ALLS_E_HIP: return "HIP runtime error";
        case ALLS_E_BUSY: return "admission refused (HBM budget)";
        case ALLS_E_NO_DEVICE: return "no GPU device";
        case ALLS_E_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

int alls_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

alls_status alls_ctx_create(alls_ctx** ctx, int device, void* hip_stream) {
    if (!ctx) return ALLS_E_INVALID;
    *ctx = nullptr;
    int n = alls_device_count();
    if (n <= 0) return ALLS_E_NO_DEVICE;
    if (device < 0 || device >= n) return ALLS_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return ALLS_E_HIP;
    alls_ctx* c = new alls_ctx();
    c->device = device;
    if (hip_stream) {
        c->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return ALLS_E_HIP;
        }
        c->own_stream = true;
    }
    if (c->ws.init() != hipSuccess) {
        c->ws.destroy();
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return ALLS_E_HIP;
    }
    // Device address of alls_assemble_noalloc's fixed scratch on this device
    // (looked up once per device; a failure is reported by that entry point).
    (void)noalloc_scratch_init(device);
    *ctx = c;
    return ALLS_OK;
}

void alls_ctx_destroy(alls_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ws.destroy();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* alls_ctx_last_error(const alls_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

alls_status alls_last_stats(const alls_ctx* ctx, alls_stats* out) {
    if (!ctx || !out) return ALLS_E_INVALID;
    *out = ctx->stats;
    return ALLS_OK;
}

alls_status alls_assemble(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts, alls_csr* out) {
    if (!ctx || !mesh || !out) return ALLS_E_INVALID;
    alls_opts o{};
    if (opts) o = *opts;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    MeshInfo mi;
    alls_status s = classify_mesh(mesh, o, &mi, &ctx->err);
    if (s != ALLS_OK) return s;
    DevMesh dm;
    s = stage_mesh(ctx->ws, mesh, mi, &dm, ctx->stream, &ctx->err);
    if (s != ALLS_OK) return s;
    s = apply_tuning(ctx, dm, mi, &o);
    if (s != ALLS_OK) return s;
    alls_stats st{};
    s = run_alls(ctx->ws, dm, mi, o, out, ctx->stream, &st, &ctx->err);
    if (s == ALLS_OK) ctx->stats = st;
    return s;
}

alls_status alls_assemble_noalloc(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts, alls_csr* out) {
    if (!ctx || !mesh || !out) return ALLS_E_INVALID;
    alls_opts o{};
    if (opts) o = *opts;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    if (mesh->mem != ALLS_MEM_DEVICE) {
        ctx->err = "alls_assemble_noalloc takes device-resident mesh arrays (host arrays need staging memory)";
        return ALLS_E_INVALID;
    }
    if (o.autotune) {
        ctx->err = "alls_assemble_noalloc does not tune (the autotuner allocates); pass explicit settings";
        return ALLS_E_INVALID;
    }
    if (!out->row_ptr || out->cap_rows < 1 || (out->cap_nnz > 0 && (!out->col_ind || !out->val))) {
        ctx->err = "alls_assemble_noalloc writes into the caller's row_ptr/col_ind/val (cap_rows, cap_nnz)";
        return ALLS_E_INVALID;
    }
    char* base = static_cast<char*>(noalloc_scratch(ctx->device));
    if (!base) {
        ctx->err = "the fixed scratch of alls_assemble_noalloc has no device address on this device";
        return ALLS_E_HIP;
    }
    MeshInfo mi;
    alls_status s = classify_mesh(mesh, o, &mi, &ctx->err);
    if (s != ALLS_OK) return s;
    // A workspace view over the fixed scratch, sharing the context's pinned
    // read-back word and events (made at context creation).
    Workspace xs;
    xs.ext = base;
    xs.ext_bytes = noalloc_scratch_bytes();
    xs.host_misc = ctx->ws.host_misc;
    for (int i = 0; i < 8; ++i) xs.ev[i] = ctx->ws.ev[i];
    xs.events = ctx->ws.events;
    alls_stats st{};
    no_alloc_scope(true);
    DevMesh dm;
    s = stage_mesh(xs, mesh, mi, &dm, ctx->stream, &ctx->err);
    if (s == ALLS_OK) s = run_alls(xs, dm, mi, o, out, ctx->stream, &st, &ctx->err);
    no_alloc_scope(false);
    // The view owns nothing: detach it before its buffers would be released.
    for (DevBuf* b : {&xs.hdr, &xs.keys, &xs.geid, &xs.vals, &xs.misc, &xs.u, &xs.big, &xs.tiles, &xs.coords,
                      &xs.conn, &xs.types, &xs.E, &xs.bdesc, &xs.bsplit}) {
        b->p = nullptr;
        b->bytes = 0;
    }
    xs.host_misc = nullptr;
    xs.events = false;
    if (s == ALLS_OK) ctx->stats = st;
    if (s == ALLS_E_OOM) {   // the allocation guard refused a buffer this path does not provide
        ctx->err = "alls_assemble_noalloc: " + ctx->err + " (device inputs must be 16-byte aligned; nothing is "
                   "allocated on this path)";
        s = ALLS_E_UNSUPPORTED;
    }
    return s;
}

alls_status alls_csr_download(alls_ctx* ctx, const alls_csr* csr, int64_t* row_ptr, int32_t* col_ind, double* val) {
    if (!ctx || !csr) return ALLS_E_INVALID;
    (void)hipSetDevice(ctx->device);
    std::string* err = &ctx->err;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    if (row_ptr && csr->row_ptr) {
        if (csr->row_ptr_is_i64) {
            HIP_TRY(hipMemcpy(row_ptr, csr->row_ptr, (size_t)(csr->n_rows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost),
                    "download row_ptr");
        } else {
            std::vector<int32_t> tmp((size_t)csr->n_rows + 1);
            HIP_TRY(hipMemcpy(tmp.data(), csr->row_ptr, tmp.size() * sizeof(int32_t), hipMemcpyDeviceToHost),
                    "download row_ptr");
            for (size_t i = 0; i < tmp.size(); ++i) row_ptr[i] = tmp[i];
        }
    }
    if (col_ind && csr->col_ind && csr->nnz)
        HIP_TRY(hipMemcpy(col_ind, csr->col_ind, (size_t)csr->nnz * sizeof(int32_t), hipMemcpyDeviceToHost),
                "download col_ind");
    if (val && csr->val && csr->nnz) {
        if (csr->val_is_f32) {
            std::vector<float> tmp((size_t)csr->nnz);
            HIP_TRY(hipMemcpy(tmp.data(), csr->val, tmp.size() * sizeof(float), hipMemcpyDeviceToHost), "download val");
            for (size_t i = 0; i < tmp.size(); ++i) val[i] = tmp[i];
        } else {
            HIP_TRY(hipMemcpy(val, csr->val, (size_t)csr->nnz * sizeof(double), hipMemcpyDeviceToHost), "download val");
        }
    }
    return ALLS_OK;
}

void alls_csr_free(alls_ctx* ctx, alls_csr* csr) {
    if (ctx) (void)hipSetDevice(ctx->device);
    free_csr(csr);
}

}  // extern "C"
