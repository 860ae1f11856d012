// pattern.cpp -- cached sparsity pattern, value-only re-assembly, GPU
// colouring and the coloured comparison path.
//
// The reference rebuilds its host pattern (Acc:605-645) for every run and
// its only value kernels are the A-LLS pair and the coloured kernels, which
// binary-search every scalar column (Acc:556-566).  A pattern object here
// keeps, on the device: the mesh geometry, row_ptr/col_ind, an element ->
// column-rank map built once by binary search and the node -> element
// incidence sorted by element id.  A request with a new E field then only
// integrates and gathers each row from its incident elements in element
// order (no lists, no atomics: k_gather_2d, or k_hex20 + k_pair_gather), or,
// for the coloured comparison, scatters with one plain launch per colour.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "host_internal.h"

using namespace alls;

struct alls_pattern {
    int device = 0;
    MeshInfo mi;
    alls_opts opts{};
    DevMesh dm{};                  // geometry in the buffers below; E is per request
    DevBuf coords, conn, types, E0;
    int64_t n_rows = 0, nnz = 0;
    int i64 = 0;
    DevBuf row_ptr, col_ind, map;
    DevBuf inc_cnt, inc_off, inc_cursor, inc_list, tiles;   // node -> element incidence, sorted by element id
    bool have_inc = false;
    int32_t umax = 0;                                       // Hex20: most column nodes in one node row
    bool repeated_node = false;                             // some element lists one node twice
    int32_t id = 0;                                         // stamped into outputs holding the pattern
};

static std::atomic<int32_t> g_pattern_ids{0};

struct alls_coloring {
    int64_t n_elems = 0;
    int32_t n_colors = 0;
    DevBuf colors, elist;
    std::vector<int64_t> offsets;   // n_colors + 1 (host copy)
};

#define HIP_TRY(call, what)                                   \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail(e_, what, err); \
    } while (0)

namespace {

alls_status copy_into(DevBuf& dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s,
                      std::string* err) {
    HIP_TRY(dst.ensure(bytes), "pattern alloc");
    if (bytes) HIP_TRY(hipMemcpyAsync(dst.p, src, bytes, kind, s), "pattern copy");
    return ALLS_OK;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

alls_status build_incidence(alls_pattern* p, hipStream_t s, std::string* err) {
    if (p->have_inc) return ALLS_OK;
    const int64_t nn = std::max<int64_t>(p->dm.n_nodes, 1);
    const int64_t nref = p->dm.n_elems * p->mi.npe_max;
    HIP_TRY(p->inc_cnt.ensure(nn * sizeof(int32_t)), "alloc incidence");
    HIP_TRY(p->inc_cursor.ensure(nn * sizeof(int32_t)), "alloc incidence");
    HIP_TRY(p->inc_off.ensure((nn + 1) * sizeof(int64_t)), "alloc incidence");
    HIP_TRY(p->inc_list.ensure(std::max<int64_t>(nref, 1) * sizeof(int32_t)), "alloc incidence");
    HIP_TRY(p->tiles.ensure((scan_tile_count(nn) + 1) * sizeof(int64_t)), "alloc incidence");
    launch_incidence(p->dm, p->inc_cnt.as<int32_t>(), p->inc_off.as<int64_t>(), p->inc_cursor.as<int32_t>(),
                     p->inc_list.as<int32_t>(), p->tiles.as<int64_t>(), s);
    HIP_TRY(hipGetLastError(), "launch incidence");
    p->have_inc = true;
    return ALLS_OK;
}

// Colour counts -> offsets -> element lists; misc is a 2-int device scratch.
alls_status bucket_colors(alls_coloring* c, const alls_pattern* p, alls_ctx* ctx, std::string* err) {
    hipStream_t s = ctx->stream;
    const int64_t n = p->dm.n_elems;
    const int32_t nc = c->n_colors;
    DevBuf cnt, cur, coff;
    HIP_TRY(cnt.ensure(std::max(nc, 1) * sizeof(int32_t)), "alloc colour counts");
    HIP_TRY(cur.ensure(std::max(nc, 1) * sizeof(int32_t)), "alloc colour counts");
    HIP_TRY(coff.ensure((nc + 1) * sizeof(int64_t)), "alloc colour offsets");
    HIP_TRY(c->elist.ensure(std::max<int64_t>(n, 1) * sizeof(int32_t)), "alloc colour lists");
    launch_fill_i32(cnt.as<int32_t>(), nc, 0, s);
    launch_fill_i32(cur.as<int32_t>(), nc, 0, s);
    launch_color_count(c->colors.as<int32_t>(), n, nc, cnt.as<int32_t>(), s);
    std::vector<int32_t> hc(nc);
    HIP_TRY(hipMemcpyAsync(hc.data(), cnt.p, nc * sizeof(int32_t), hipMemcpyDeviceToHost, s), "read colour counts");
    HIP_TRY(hipStreamSynchronize(s), "sync colour counts");
    c->offsets.assign(nc + 1, 0);
    for (int32_t i = 0; i < nc; ++i) c->offsets[i + 1] = c->offsets[i] + hc[i];
    HIP_TRY(hipMemcpyAsync(coff.p, c->offsets.data(), (nc + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s),
            "write colour offsets");
    launch_color_bucket(c->colors.as<int32_t>(), n, nc, coff.as<int64_t>(), cur.as<int32_t>(), c->elist.as<int32_t>(),
                        s);
    HIP_TRY(hipStreamSynchronize(s), "sync colour buckets");
    cnt.release();
    cur.release();
    coff.release();
    return ALLS_OK;
}

// E field of a request: device pointer, staged host copy, or the pattern's own.
alls_status stage_E(alls_ctx* ctx, const alls_pattern* p, const double* E, int32_t mem, const double** out,
                    std::string* err) {
    if (!E) {
        *out = p->dm.E_elem;
        return ALLS_OK;
    }
    if (mem == ALLS_MEM_DEVICE) {
        *out = E;
        return ALLS_OK;
    }
    const size_t eb = (size_t)p->dm.n_elems * sizeof(double);
    HIP_TRY(ctx->ws.E.ensure(std::max<size_t>(eb, 8)), "alloc E");
    void* dst[1] = {ctx->ws.E.p};
    const void* src[1] = {E};
    HIP_TRY(ctx->ws.upload(dst, src, &eb, 1, ctx->stream), "upload E");
    *out = ctx->ws.E.as<double>();
    return ALLS_OK;
}

// Shares the pattern's row_ptr/col_ind with `out` (copied once per buffer).
alls_status pattern_into(alls_pattern* p, alls_csr* out, hipStream_t s, std::string* err) {
    alls_status st = ensure_csr(out, p->n_rows, p->nnz, p->i64);
    if (st != ALLS_OK) {
        if (err) *err = "cannot allocate output CSR";
        return st;
    }
    if (out->pattern_id != p->id) {
        HIP_TRY(hipMemcpyAsync(out->row_ptr, p->row_ptr.p, (p->n_rows + 1) * (p->i64 ? 8 : 4),
                               hipMemcpyDeviceToDevice, s), "copy row_ptr");
        if (p->nnz)
            HIP_TRY(hipMemcpyAsync(out->col_ind, p->col_ind.p, p->nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s),
                    "copy col_ind");
        out->pattern_id = p->id;
    }
    return ALLS_OK;
}

Scatter make_scatter(const alls_pattern* p, alls_csr* out) {
    Scatter sc{};
    sc.map = p->map.as<int32_t>();
    sc.row_ptr = p->row_ptr.p;
    sc.val = out->val;
    sc.i64 = p->i64;
    sc.npe = p->mi.kind == 3 ? 20 : 4;
    return sc;
}

void launch_scatter(const alls_pattern* p, const DevMesh& dm, const Scatter& sc, int atomic, int32_t* misc,
                    hipStream_t s) {
    if (p->mi.kind == 3) launch_scatter_hex20(dm, sc, atomic, p->mi.f32 ? 1 : 0, misc, s);
    else launch_scatter_2d(dm, p->mi.kind, sc, atomic, misc, s);
}

// Hex20: the largest row (column nodes) sizes the ordered gather's LDS.
alls_status hex20_row_limit(alls_pattern* p, int32_t* misc, hipStream_t s, std::string* err) {
    launch_fill_i32(misc, MISC_SIZE, 0, s);
    launch_row_umax(p->dm.n_nodes, p->row_ptr.p, p->i64, misc, s);
    HIP_TRY(hipGetLastError(), "launch row size");
    HIP_TRY(hipMemcpyAsync(&p->umax, misc, sizeof(int32_t), hipMemcpyDeviceToHost, s), "read row size");
    HIP_TRY(hipStreamSynchronize(s), "sync row size");
    return ALLS_OK;
}

}  // namespace

namespace alls {

// Used by the server: pattern from an already-staged mesh on stream s.
alls_status pattern_build(Workspace& ws, const alls_mesh* mesh, const alls_opts& o, hipStream_t s, int device,
                          alls_pattern** out_pat, alls_csr* out, alls_stats* st, std::string* err) {
    MeshInfo mi;
    alls_status rc = classify_mesh(mesh, o, &mi, err);
    if (rc != ALLS_OK) return rc;
    DevMesh staged;
    rc = stage_mesh(ws, mesh, mi, &staged, s, err);
    if (rc != ALLS_OK) return rc;
    alls_pattern* p = new alls_pattern();
    p->device = device;
    p->id = 0;
    p->mi = mi;
    p->opts = o;
    p->dm = staged;
    const size_t cb = (size_t)mesh->n_nodes * mesh->dim * sizeof(double);
    const size_t kb = (size_t)mesh->n_elems * mesh->nodes_per_elem * sizeof(int32_t);
    auto fail = [&](alls_status r) {
        delete p;
        return r;
    };
    if ((rc = copy_into(p->coords, staged.coords, cb, hipMemcpyDeviceToDevice, s, err)) != ALLS_OK) return fail(rc);
    if ((rc = copy_into(p->conn, staged.conn, kb, hipMemcpyDeviceToDevice, s, err)) != ALLS_OK) return fail(rc);
    p->dm.coords = p->coords.as<double>();
    p->dm.conn = p->conn.as<int32_t>();
    if (staged.types) {
        if ((rc = copy_into(p->types, staged.types, (size_t)mesh->n_elems, hipMemcpyDeviceToDevice, s, err)) != ALLS_OK)
            return fail(rc);
        p->dm.types = p->types.as<uint8_t>();
    }
    if (staged.E_elem) {
        if ((rc = copy_into(p->E0, staged.E_elem, (size_t)mesh->n_elems * sizeof(double), hipMemcpyDeviceToDevice, s,
                            err)) != ALLS_OK)
            return fail(rc);
        p->dm.E_elem = p->E0.as<double>();
    }
    rc = run_alls(ws, p->dm, mi, o, out, s, st, err);
    if (rc != ALLS_OK) return fail(rc);
    p->n_rows = out->n_rows;
    p->nnz = out->nnz;
    p->i64 = out->row_ptr_is_i64;
    if ((rc = copy_into(p->row_ptr, out->row_ptr, (p->n_rows + 1) * (p->i64 ? 8 : 4), hipMemcpyDeviceToDevice, s,
                        err)) != ALLS_OK)
        return fail(rc);
    if ((rc = copy_into(p->col_ind, out->col_ind, std::max<int64_t>(p->nnz, 1) * sizeof(int32_t),
                        hipMemcpyDeviceToDevice, s, err)) != ALLS_OK)
        return fail(rc);
    p->id = ++g_pattern_ids;
    out->pattern_id = p->id;
    const int npm = mi.kind == 3 ? 20 : 4;
    if (p->map.ensure(std::max<int64_t>(mesh->n_elems, 1) * npm * npm * sizeof(int32_t)) != hipSuccess)
        return fail(hip_fail(hipErrorOutOfMemory, "alloc scatter map", err));
    if ((rc = build_incidence(p, s, err)) != ALLS_OK) return fail(rc);
    launch_fill_i32(ws.misc.as<int32_t>(), MISC_SIZE, 0, s);
    launch_build_map(p->dm, mi.kind, p->row_ptr.p, p->i64, p->col_ind.as<int32_t>(), p->inc_off.as<int64_t>(),
                     p->inc_list.as<int32_t>(), p->map.as<int32_t>(), ws.misc.as<int32_t>(), s);
    int32_t flags[MISC_SIZE];
    if (hipMemcpyAsync(flags, ws.misc.p, sizeof(flags), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "build scatter map", err));
    if (flags[MISC_ERR + ERR_INVALID_MESH]) {
        if (err) *err = "scatter map: a column was missing from the pattern";
        return fail(ALLS_E_HIP);
    }
    p->repeated_node = flags[MISC_ERR + ERR_REPEATED_NODE] != 0;
    if (mi.kind == 3 && (rc = hex20_row_limit(p, ws.misc.as<int32_t>(), s, err)) != ALLS_OK) return fail(rc);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(hip_fail(hipGetLastError(), "sync pattern", err));
    *out_pat = p;
    return ALLS_OK;
}

// Value-only assembly on stream s with workspace scratch for flags/events.
alls_status pattern_values(Workspace& ws, alls_pattern* p, const double* E_dev, double E_scalar, alls_csr* out,
                           hipStream_t s, alls_stats* st, std::string* err) {
    HIP_TRY(ws.init(), "workspace init");
    alls_status rc = pattern_into(p, out, s, err);
    if (rc != ALLS_OK) return rc;
    HIP_TRY(ws.misc.ensure(MISC_SIZE * sizeof(int32_t)), "alloc misc");
    DevMesh dm = p->dm;
    dm.E_elem = E_dev;
    dm.E = E_scalar;
    // Hex20 rows too wide for the gather's LDS, or an element that lists a
    // node twice (two of its blocks would land on one column in the same
    // gather step): the plain pipeline, whose values are the same bits.
    if (p->mi.kind == 3 && (p->umax > pair_gather_max_cols() || p->repeated_node))
        return run_alls(ws, dm, p->mi, p->opts, out, s, st, err);
    Scatter sc = make_scatter(p, out);
    sc.count = dm.n_elems;
    HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
    launch_fill_i32(ws.misc.as<int32_t>(), MISC_SIZE, 0, s);
    if (p->mi.kind == 3) {
        // Integrate into element-order staging, then gather each node row in
        // element order.
        HIP_TRY(ws.vals.ensure((size_t)std::max<int64_t>(dm.n_elems, 1) * 400 * 9 * (p->mi.f32 ? 4 : 8)),
                "alloc stage");
        sc.stage = ws.vals.p;
        HIP_TRY(hipEventRecord(ws.ev[1], s), "event");
        launch_scatter_hex20(dm, sc, 3, p->mi.f32 ? 1 : 0, ws.misc.as<int32_t>(), s);
        HIP_TRY(hipEventRecord(ws.ev[2], s), "event");
        launch_pair_gather(dm.n_nodes, dm.conn, p->row_ptr.p, p->i64, p->map.as<int32_t>(), p->inc_off.as<int64_t>(),
                           p->inc_list.as<int32_t>(), ws.vals.p, p->mi.f32 ? 1 : 0, p->umax, out->val, s);
    } else {
        HIP_TRY(hipEventRecord(ws.ev[1], s), "event");
        HIP_TRY(hipEventRecord(ws.ev[2], s), "event");
        launch_gather_2d(dm, p->mi.kind, sc, p->inc_off.as<int64_t>(), p->inc_list.as<int32_t>(),
                         ws.misc.as<int32_t>(), s);
    }
    HIP_TRY(hipGetLastError(), "launch value-only");
    HIP_TRY(hipEventRecord(ws.ev[3], s), "event");
    HIP_TRY(hipStreamSynchronize(s), "sync values");
    if (st) {
        *st = alls_stats{};
        st->ms_init = ev_ms(ws.ev[0], ws.ev[1]);
        st->ms_insert = ev_ms(ws.ev[1], ws.ev[2]);
        st->ms_write = ev_ms(ws.ev[2], ws.ev[3]);
        st->ms_total = ev_ms(ws.ev[0], ws.ev[3]);
        st->n_elems = dm.n_elems;
        st->n_rows = p->n_rows;
        st->nnz = p->nnz;
    }
    return ALLS_OK;
}

void pattern_destroy(alls_pattern* p) { delete p; }
int64_t pattern_bytes(const alls_pattern* p) {
    return (int64_t)(p->coords.bytes + p->conn.bytes + p->types.bytes + p->E0.bytes + p->row_ptr.bytes +
                     p->col_ind.bytes + p->map.bytes + p->inc_cnt.bytes + p->inc_off.bytes + p->inc_cursor.bytes +
                     p->inc_list.bytes + p->tiles.bytes);
}

}  // namespace alls

extern "C" {

alls_status alls_pattern_create(alls_ctx* ctx, const alls_mesh* mesh, const alls_opts* opts, alls_pattern** pat,
                                alls_csr* out) {
    if (!ctx || !mesh || !pat || !out) return ALLS_E_INVALID;
    *pat = nullptr;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    alls_opts o{};
    if (opts) o = *opts;
    if (o.val_f32) {
        ctx->err = "FP32 output values (opts.val_f32) are for alls_assemble and alls_assemble_noalloc";
        return ALLS_E_UNSUPPORTED;
    }
    alls_stats st{};
    alls_status rc = pattern_build(ctx->ws, mesh, o, ctx->stream, ctx->device, pat, out, &st, &ctx->err);
    if (rc == ALLS_OK) ctx->stats = st;
    return rc;
}

alls_status alls_assemble_values(alls_ctx* ctx, const alls_pattern* cpat, const double* E_per_elem, int32_t mem,
                                 alls_csr* out) {
    if (!ctx || !cpat || !out) return ALLS_E_INVALID;
    alls_pattern* p = const_cast<alls_pattern*>(cpat);
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    const double* E = nullptr;
    alls_status rc = stage_E(ctx, p, E_per_elem, mem, &E, &ctx->err);
    if (rc != ALLS_OK) return rc;
    alls_stats st{};
    rc = pattern_values(ctx->ws, p, E, p->dm.E, out, ctx->stream, &st, &ctx->err);
    if (rc == ALLS_OK) ctx->stats = st;
    return rc;
}

int64_t alls_pattern_nnz(const alls_pattern* p) { return p ? p->nnz : -1; }

void alls_pattern_free(alls_ctx* ctx, alls_pattern* p) {
    if (!p) return;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    delete p;
}

alls_status alls_color(alls_ctx* ctx, const alls_pattern* cpat, alls_coloring** col) {
    if (!ctx || !cpat || !col) return ALLS_E_INVALID;
    *col = nullptr;
    alls_pattern* p = const_cast<alls_pattern*>(cpat);
    std::string* err = &ctx->err;
    err->clear();
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    alls_status rc = build_incidence(p, s, err);
    if (rc != ALLS_OK) return rc;
    const int64_t n = p->dm.n_elems;
    alls_coloring* c = new alls_coloring();
    c->n_elems = n;
    DevBuf other;
    if (c->colors.ensure(std::max<int64_t>(n, 1) * sizeof(int32_t)) != hipSuccess ||
        other.ensure(std::max<int64_t>(n, 1) * sizeof(int32_t)) != hipSuccess ||
        ctx->ws.misc.ensure(MISC_SIZE * sizeof(int32_t)) != hipSuccess) {
        delete c;
        return hip_fail(hipErrorOutOfMemory, "alloc colours", err);
    }
    launch_fill_i32(c->colors.as<int32_t>(), n, -1, s);
    int32_t* cin = c->colors.as<int32_t>();
    int32_t* cout = other.as<int32_t>();
    int32_t h[2] = {0, 0};
    int rounds = 0;
    for (; n > 0 && rounds < 100000; ++rounds) {
        launch_fill_i32(ctx->ws.misc.as<int32_t>(), 2, 0, s);
        launch_jpl_round(p->dm, p->inc_off.as<int64_t>(), p->inc_list.as<int32_t>(), cin, cout,
                         ctx->ws.misc.as<int32_t>(), s);
        std::swap(cin, cout);
        if (hipMemcpyAsync(h, ctx->ws.misc.p, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            delete c;
            return hip_fail(hipGetLastError(), "colour round", err);
        }
        if (h[0] == 0) break;
    }
    if (cin != c->colors.as<int32_t>()) std::swap(c->colors, other);
    // Largest colour: rerun the conflict check, which also reports it.
    launch_fill_i32(ctx->ws.misc.as<int32_t>(), 2, 0, s);
    launch_conflicts(p->dm, p->inc_off.as<int64_t>(), p->inc_list.as<int32_t>(), c->colors.as<int32_t>(),
                     ctx->ws.misc.as<int32_t>(), s);
    if (hipMemcpyAsync(h, ctx->ws.misc.p, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        delete c;
        return hip_fail(hipGetLastError(), "colour check", err);
    }
    if (h[0] != 0) {
        *err = "GPU colouring produced a conflict";
        delete c;
        return ALLS_E_HIP;
    }
    c->n_colors = n > 0 ? h[1] + 1 : 0;
    rc = bucket_colors(c, p, ctx, err);
    if (rc != ALLS_OK) {
        delete c;
        return rc;
    }
    ctx->stats.n_colors = c->n_colors;
    *col = c;
    return ALLS_OK;
}

alls_status alls_coloring_from_host(alls_ctx* ctx, const alls_pattern* cpat, const int32_t* colors,
                                    alls_coloring** col) {
    if (!ctx || !cpat || !colors || !col) return ALLS_E_INVALID;
    *col = nullptr;
    alls_pattern* p = const_cast<alls_pattern*>(cpat);
    std::string* err = &ctx->err;
    err->clear();
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    alls_status rc = build_incidence(p, s, err);
    if (rc != ALLS_OK) return rc;
    const int64_t n = p->dm.n_elems;
    alls_coloring* c = new alls_coloring();
    c->n_elems = n;
    if (c->colors.ensure(std::max<int64_t>(n, 1) * sizeof(int32_t)) != hipSuccess ||
        ctx->ws.misc.ensure(MISC_SIZE * sizeof(int32_t)) != hipSuccess) {
        delete c;
        return hip_fail(hipErrorOutOfMemory, "alloc colours", err);
    }
    int32_t h[2] = {0, 0};
    launch_fill_i32(ctx->ws.misc.as<int32_t>(), 2, 0, s);
    if (hipMemcpyAsync(c->colors.p, colors, n * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess) {
        delete c;
        return hip_fail(hipGetLastError(), "upload colours", err);
    }
    launch_conflicts(p->dm, p->inc_off.as<int64_t>(), p->inc_list.as<int32_t>(), c->colors.as<int32_t>(),
                     ctx->ws.misc.as<int32_t>(), s);
    if (hipMemcpyAsync(h, ctx->ws.misc.p, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        delete c;
        return hip_fail(hipGetLastError(), "colour check", err);
    }
    if (h[0] != 0) {
        *err = "colouring has adjacent elements of equal colour (or a negative colour)";
        delete c;
        return ALLS_E_INVALID;
    }
    c->n_colors = n > 0 ? h[1] + 1 : 0;
    rc = bucket_colors(c, p, ctx, err);
    if (rc != ALLS_OK) {
        delete c;
        return rc;
    }
    *col = c;
    return ALLS_OK;
}

int32_t alls_coloring_count(const alls_coloring* c) { return c ? c->n_colors : -1; }

alls_status alls_coloring_download(alls_ctx* ctx, const alls_coloring* c, int32_t* colors) {
    if (!ctx || !c || !colors) return ALLS_E_INVALID;
    (void)hipSetDevice(ctx->device);
    std::string* err = &ctx->err;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    if (c->n_elems)
        HIP_TRY(hipMemcpy(colors, c->colors.p, c->n_elems * sizeof(int32_t), hipMemcpyDeviceToHost), "download colours");
    return ALLS_OK;
}

void alls_coloring_free(alls_ctx* ctx, alls_coloring* c) {
    if (!c) return;
    if (ctx) (void)hipSetDevice(ctx->device);
    delete c;
}

alls_status alls_assemble_colored(alls_ctx* ctx, const alls_pattern* cpat, const alls_coloring* col,
                                  const double* E_per_elem, int32_t mem, alls_csr* out) {
    if (!ctx || !cpat || !col || !out) return ALLS_E_INVALID;
    alls_pattern* p = const_cast<alls_pattern*>(cpat);
    std::string* err = &ctx->err;
    err->clear();
    if (col->n_elems != p->dm.n_elems) {
        *err = "colouring belongs to another mesh";
        return ALLS_E_INVALID;
    }
    if (p->mi.kind == 3 && p->repeated_node) {
        // The Hex20 coloured kernel adds an element's 400 blocks from
        // parallel lanes: two blocks of one element on one column would race.
        *err = "coloured Hex20 assembly of an element that lists a node twice";
        return ALLS_E_UNSUPPORTED;
    }
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    Workspace& ws = ctx->ws;
    HIP_TRY(ws.init(), "workspace init");
    const double* E = nullptr;
    alls_status rc = stage_E(ctx, p, E_per_elem, mem, &E, err);
    if (rc != ALLS_OK) return rc;
    rc = pattern_into(p, out, s, err);
    if (rc != ALLS_OK) return rc;
    HIP_TRY(ws.misc.ensure(MISC_SIZE * sizeof(int32_t)), "alloc misc");
    DevMesh dm = p->dm;
    dm.E_elem = E;
    HIP_TRY(hipEventRecord(ws.ev[0], s), "event");
    launch_fill_i32(ws.misc.as<int32_t>(), MISC_SIZE, 0, s);
    launch_fill_f64(out->val, p->nnz, 0.0, s);
    HIP_TRY(hipEventRecord(ws.ev[1], s), "event");
    // One full-grid launch per colour: no two elements of a launch share a node.
    for (int32_t c = 0; c < col->n_colors; ++c) {
        Scatter sc = make_scatter(p, out);
        sc.elist = col->elist.as<int32_t>() + col->offsets[c];
        sc.count = col->offsets[c + 1] - col->offsets[c];
        launch_scatter(p, dm, sc, 0, ws.misc.as<int32_t>(), s);
    }
    HIP_TRY(hipGetLastError(), "launch coloured scatter");
    HIP_TRY(hipEventRecord(ws.ev[2], s), "event");
    HIP_TRY(hipStreamSynchronize(s), "sync coloured");
    alls_stats st{};
    st.ms_init = ev_ms(ws.ev[0], ws.ev[1]);
    st.ms_insert = ev_ms(ws.ev[1], ws.ev[2]);
    st.ms_total = ev_ms(ws.ev[0], ws.ev[2]);
    st.n_elems = dm.n_elems;
    st.n_rows = p->n_rows;
    st.nnz = p->nnz;
    st.n_colors = col->n_colors;
    ctx->stats = st;
    return ALLS_OK;
}

}  // extern "C"
