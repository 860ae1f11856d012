// batch.cpp -- many meshes in one pipeline, and the launch autotuner.
//
// Batching: the meshes of one dimension are packed into one node/element
// space (mesh k's node ids shifted by the nodes of meshes 0..k-1), so every
// stage of the A-LLS pipeline is ONE launch over all of them, whatever their
// sizes, and the result is the block-diagonal CSR of the batch.  The
// insertion kernel runs persistent (a grid sized to the device, looping over
// the packed elements) and the whole pipeline is captured as ONE hipGraph
// (its output sized by the list entries bound, so no host read-back splits
// it) and replayed while the batch keeps its shape.  The reference has no batched path: each of its programs
// assembles one mesh per call (Acc:750-838, T3:525, Mix:578, Hex:420).
//
// Autotuning: block size, persistent grid and list-pool chunk size are
// measured on the mesh itself (device time of the whole pipeline) and the
// fastest is kept per (element kind, precision, size class).  Results are
// bit-identical across settings: consolidation sums every column in element
// order whatever the list layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"

namespace alls {

static alls_status fail(std::string* err, alls_status s, const std::string& msg) {
    if (err) *err = msg;
    return s;
}

#define HIP_TRY(call, what)                                   \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail(e_, what, err); \
    } while (0)

static uint64_t tune_key(const MeshInfo& mi, int64_t n_elems) {
    int sz = 0;
    while (sz < 62 && (int64_t(1) << (sz + 1)) <= std::max<int64_t>(n_elems, 1)) ++sz;
    return ((uint64_t)mi.kind << 16) | ((uint64_t)(mi.f32 ? 1 : 0) << 8) | (uint64_t)sz;
}

// Measures every candidate on dm (two runs each, the faster counts).
static alls_status tune(alls_ctx* ctx, const DevMesh& dm, const MeshInfo& mi, const alls_opts& base, alls_tuned* best) {
    std::vector<int> blocks, grids, chunks;
    if (mi.kind == 3) {
        blocks = {64};
        grids = {0, 16, 32};
        chunks = {40, 80, 160};
    } else {
        blocks = {64, 128, 256};
        grids = {0, 8, 32};
        chunks = mi.kind == 0 ? std::vector<int>{8, 16, 32} : std::vector<int>{12, 24, 48};
    }
    alls_csr tmp{};
    best->ms = -1.0;
    for (int b : blocks)
        for (int g : grids)
            for (int c : chunks) {
                alls_opts o = base;
                o.block_threads = b;
                o.grid_per_cu = g;
                o.chunk_slots = c;
                o.autotune = 0;
                o.timing = 1;   // candidates are ranked by their device time
                double ms = 1e300;
                for (int rep = 0; rep < 2; ++rep) {
                    alls_stats st{};
                    const alls_status s = run_alls(ctx->ws, dm, mi, o, &tmp, ctx->stream, &st, &ctx->err);
                    if (s != ALLS_OK) {
                        free_csr(&tmp);
                        return s;
                    }
                    ms = std::min(ms, st.ms_total);
                }
                if (best->ms < 0 || ms < best->ms) *best = alls_tuned{b, c, g, ms};
            }
    free_csr(&tmp);
    return ALLS_OK;
}

// opts.autotune: fills block_threads / chunk_slots / grid_per_cu from the
// context's tuned table, tuning on dm when this kind and size class is new.
alls_status apply_tuning(alls_ctx* ctx, const DevMesh& dm, const MeshInfo& mi, alls_opts* o) {
    if (!o->autotune) return ALLS_OK;
    const uint64_t k = tune_key(mi, dm.n_elems);
    auto it = ctx->tuned.find(k);
    if (it == ctx->tuned.end()) {
        alls_tuned t{};
        const alls_status s = tune(ctx, dm, mi, *o, &t);
        if (s != ALLS_OK) return s;
        it = ctx->tuned.emplace(k, t).first;
    }
    o->block_threads = it->second.block_threads;
    o->chunk_slots = it->second.chunk_slots;
    o->grid_per_cu = it->second.grid_per_cu;
    return ALLS_OK;
}

}  // namespace alls

using namespace alls;

extern "C" {

alls_status alls_autotune(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* base, alls_opts* best) {
    if (!ctx || !mesh || !best) return ALLS_E_INVALID;
    alls_opts o{};
    if (base) o = *base;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    MeshInfo mi;
    alls_status s = classify_mesh(mesh, o, &mi, &ctx->err);
    if (s != ALLS_OK) return s;
    DevMesh dm;
    s = stage_mesh(ctx->ws, mesh, mi, &dm, ctx->stream, &ctx->err);
    if (s != ALLS_OK) return s;
    alls_tuned t{};
    s = tune(ctx, dm, mi, o, &t);
    if (s != ALLS_OK) return s;
    ctx->tuned[tune_key(mi, dm.n_elems)] = t;
    *best = o;
    best->block_threads = t.block_threads;
    best->chunk_slots = t.chunk_slots;
    best->grid_per_cu = t.grid_per_cu;
    best->autotune = 0;
    return ALLS_OK;
}

alls_status alls_assemble_batch(alls_ctx* ctx, int32_t n, const alls_mesh* meshes, const alls_opts* opts,
                                alls_csr* out, int64_t* row_off, int64_t* nnz_off) {
    if (!ctx || n < 1 || !meshes || !out || !row_off || !nnz_off) return ALLS_E_INVALID;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    alls_opts o{};
    if (opts) o = *opts;
    if (o.val_f32) {
        ctx->err = "FP32 output values (opts.val_f32) are for alls_assemble and alls_assemble_noalloc";
        return ALLS_E_UNSUPPORTED;
    }
    alls_stats stats{};
    const alls_status st = batch_assemble(ctx->ws, ctx->stream, n, meshes, o, out, row_off, nnz_off, &stats,
                                          &ctx->err, ctx, /*use_graph=*/true);
    if (st == ALLS_OK) ctx->stats = stats;
    return st;
}

}  // extern "C"

namespace alls {

alls_status batch_assemble(Workspace& ws, hipStream_t s, int32_t n, const alls_mesh* meshes, alls_opts o,
                           alls_csr* out, int64_t* row_off, int64_t* nnz_off, alls_stats* stats_out,
                           std::string* err, alls_ctx* tune_ctx, bool use_graph) {

    // Validate and lay out the packed mesh.
    std::vector<MeshInfo> mis((size_t)n);
    std::vector<int64_t> node_off((size_t)n + 1, 0), elem_off((size_t)n + 1, 0);
    bool any_q4 = false, any_t3 = false, mixed = false, need_E = false;
    int64_t entries = 0;
    for (int k = 0; k < n; ++k) {
        const alls_mesh& m = meshes[k];
        alls_status st = classify_mesh(&m, o, &mis[k], err);
        if (st != ALLS_OK) return fail(err, st, "mesh " + std::to_string(k) + ": " + *err);
        if (m.dim != meshes[0].dim || m.mem != meshes[0].mem)
            return fail(err, ALLS_E_INVALID, "meshes of a batch must share dim and memory space");
        if (m.nu != meshes[0].nu || (m.dim == 2 && m.thickness != meshes[0].thickness))
            return fail(err, ALLS_E_INVALID, "meshes of a batch must share nu and thickness");
        any_q4 |= mis[k].kind == 0;
        any_t3 |= mis[k].kind == 1;
        mixed |= mis[k].kind == 2;
        need_E |= m.E_per_elem != nullptr || m.E != meshes[0].E;
        node_off[k + 1] = node_off[k] + m.n_nodes;
        elem_off[k + 1] = elem_off[k] + m.n_elems;
        entries += mis[k].entries_bound;
    }
    const int dim = meshes[0].dim;
    const int64_t N = node_off[n], Ne = elem_off[n];
    if (N >= (int64_t)INT32_MAX / 3 || Ne >= (int64_t)INT32_MAX)
        return fail(err, ALLS_E_INVALID, "batch too large for 32-bit node/element ids");
    MeshInfo mi = mis[0];
    int uniform = ALLS_HEX20;
    if (dim == 2) {
        if (mixed || (any_q4 && any_t3)) {
            mi.kind = 2;
            mi.npe_max = 4;
        } else if (any_t3) {
            mi.kind = 1;
            mi.npe_max = 3;
            uniform = ALLS_T3;
        } else {
            mi.kind = 0;
            mi.npe_max = 4;
            uniform = ALLS_Q4;
        }
    }
    mi.entries_bound = entries;
    const int npe = dim == 3 ? 20 : (mi.kind == 1 ? 3 : 4);
    const bool with_types = mi.kind == 2;

    const size_t cb = (size_t)N * dim * sizeof(double);
    const size_t kb = (size_t)Ne * npe * sizeof(int32_t);
    const size_t tb = with_types ? (size_t)Ne : 0;
    const size_t eb = need_E ? (size_t)Ne * sizeof(double) : 0;
    HIP_TRY(ws.init(), "workspace init");
    HIP_TRY(ws.coords.ensure(std::max<size_t>(cb, 16)), "alloc coords");
    HIP_TRY(ws.conn.ensure(std::max<size_t>(kb, 16)), "alloc conn");
    if (tb) HIP_TRY(ws.types.ensure(tb), "alloc types");
    if (eb) HIP_TRY(ws.E.ensure(eb), "alloc E");
    const size_t desc_bytes = (size_t)n * sizeof(PackMesh);
    const size_t idx_off = (desc_bytes + 255) & ~size_t(255);
    const size_t res_off = idx_off + (((size_t)(n + 1) * sizeof(int64_t) + 255) & ~size_t(255));
    HIP_TRY(ws.bdesc.ensure(res_off + (size_t)(n + 1) * sizeof(int64_t)), "alloc batch table");

    if (meshes[0].mem == ALLS_MEM_HOST) {
        // Host meshes: packed straight into the pinned buffer, four copies.
        const size_t o1 = (cb + 255) & ~size_t(255), o2 = o1 + ((kb + 255) & ~size_t(255));
        const size_t o3 = o2 + ((tb + 255) & ~size_t(255)), total = o3 + eb;
        char* h = nullptr;
        HIP_TRY(ws.staging(std::max<size_t>(total, 256), &h), "pinned staging");
        double* hc = reinterpret_cast<double*>(h);
        int32_t* hk = reinterpret_cast<int32_t*>(h + o1);
        uint8_t* ht = reinterpret_cast<uint8_t*>(h + o2);
        double* hE = reinterpret_cast<double*>(h + o3);
        for (int k = 0; k < n; ++k) {
            const alls_mesh& m = meshes[k];
            if (m.n_nodes) std::memcpy(hc + node_off[k] * dim, m.coords, (size_t)m.n_nodes * dim * sizeof(double));
            const int32_t sh = (int32_t)node_off[k];
            for (int64_t e = 0; e < m.n_elems; ++e) {
                const int32_t* src = m.conn + e * m.nodes_per_elem;
                int32_t* dst = hk + (elem_off[k] + e) * npe;
                for (int j = 0; j < npe; ++j) {
                    const int32_t c = j < m.nodes_per_elem ? src[j] : -1;
                    dst[j] = c >= 0 ? c + sh : -1;
                }
            }
            if (with_types) {
                if (m.elem_type) std::memcpy(ht + elem_off[k], m.elem_type, (size_t)m.n_elems);
                else std::memset(ht + elem_off[k], m.uniform_type, (size_t)m.n_elems);
            }
            if (need_E) {
                if (m.E_per_elem) std::memcpy(hE + elem_off[k], m.E_per_elem, (size_t)m.n_elems * sizeof(double));
                else std::fill(hE + elem_off[k], hE + elem_off[k] + m.n_elems, m.E);
            }
        }
        if (cb) HIP_TRY(hipMemcpyAsync(ws.coords.p, hc, cb, hipMemcpyHostToDevice, s), "upload coords");
        if (kb) HIP_TRY(hipMemcpyAsync(ws.conn.p, hk, kb, hipMemcpyHostToDevice, s), "upload conn");
        if (tb) HIP_TRY(hipMemcpyAsync(ws.types.p, ht, tb, hipMemcpyHostToDevice, s), "upload types");
        if (eb) HIP_TRY(hipMemcpyAsync(ws.E.p, hE, eb, hipMemcpyHostToDevice, s), "upload E");
    } else {
        // Device meshes: a table of sources, then one packing launch each
        // for nodes and elements.
        char* h = nullptr;
        HIP_TRY(ws.staging(std::max<size_t>(desc_bytes, 256), &h), "pinned staging");
        PackMesh* pm = reinterpret_cast<PackMesh*>(h);
        for (int k = 0; k < n; ++k) {
            const alls_mesh& m = meshes[k];
            pm[k] = PackMesh{m.coords, m.conn, m.elem_type, m.E_per_elem, m.E, m.n_nodes, m.n_elems,
                             node_off[k], elem_off[k], m.nodes_per_elem, m.uniform_type};
        }
        HIP_TRY(hipMemcpyAsync(ws.bdesc.p, pm, desc_bytes, hipMemcpyHostToDevice, s), "upload batch table");
        launch_pack(ws.bdesc.as<PackMesh>(), n, dim, npe, N, Ne, ws.coords.as<double>(), ws.conn.as<int32_t>(),
                    with_types ? ws.types.as<uint8_t>() : nullptr, need_E ? ws.E.as<double>() : nullptr, s);
        HIP_TRY(hipGetLastError(), "launch pack");
    }

    DevMesh dm{};
    dm.coords = ws.coords.as<double>();
    dm.conn = ws.conn.as<int32_t>();
    dm.types = with_types ? ws.types.as<uint8_t>() : nullptr;
    dm.E_elem = need_E ? ws.E.as<double>() : nullptr;
    dm.E = meshes[0].E;
    dm.nu = meshes[0].nu;
    dm.thickness = meshes[0].thickness;
    dm.n_nodes = N;
    dm.n_elems = Ne;
    dm.dim = dim;
    dm.npe = npe;
    dm.uniform_type = uniform;
    // Persistent by default: 32 blocks per CU for Hex20 (a 110k-element batch
    // of 8 meshes: FP64 3.60 -> 2.64 ms, FP32 2.73 -> 2.05 ms against 4 per
    // CU; 64^3: 8.56 -> 6.25 / 6.36 -> 4.74 ms), 8 for 2D.
    if (o.grid_per_cu <= 0 && !o.autotune) o.grid_per_cu = dim == 3 ? 32 : 8;
    alls_status st = ALLS_OK;
    if (tune_ctx) st = apply_tuning(tune_ctx, dm, mi, &o);
    if (st != ALLS_OK) return st;
    alls_stats stats{};
    st = run_alls(ws, dm, mi, o, out, s, &stats, err, use_graph && N > 0);
    if (st != ALLS_OK) return st;

    // Offsets of each mesh's rows and nonzeros in the block-diagonal result.
    char* h = nullptr;
    HIP_TRY(ws.staging(2 * (size_t)(n + 1) * sizeof(int64_t), &h), "pinned staging");
    int64_t* hidx = reinterpret_cast<int64_t*>(h);
    int64_t* hres = hidx + (n + 1);
    for (int k = 0; k <= n; ++k) hidx[k] = row_off[k] = node_off[k] * dim;
    char* dbase = static_cast<char*>(ws.bdesc.p);
    HIP_TRY(hipMemcpyAsync(dbase + idx_off, hidx, (size_t)(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s),
            "upload boundaries");
    launch_gather_rp(out->row_ptr, out->row_ptr_is_i64, reinterpret_cast<int64_t*>(dbase + idx_off), n + 1,
                     reinterpret_cast<int64_t*>(dbase + res_off), s);
    HIP_TRY(hipGetLastError(), "launch gather");
    HIP_TRY(hipMemcpyAsync(hres, dbase + res_off, (size_t)(n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s),
            "read offsets");
    HIP_TRY(hipStreamSynchronize(s), "sync");
    std::memcpy(nnz_off, hres, (size_t)(n + 1) * sizeof(int64_t));
    stats.n_elems = Ne;
    if (stats_out) *stats_out = stats;
    return ALLS_OK;
}

}  // namespace alls
