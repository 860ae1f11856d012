// io.cpp -- CSR and mesh files.
//
// alls_csr_write_mtx writes an assembled matrix as a Matrix Market
// coordinate file, the exchange format sparse solvers read; the reference
// writes only CSV timing tables (Perf:1139, Mix:838).  alls_mesh_save /
// alls_mesh_load keep a mesh in a small binary file (a fixed header, then
// the structure-of-arrays arrays as they are in memory), so benchmark and
// served meshes can be stored and reloaded without regenerating them.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"

using namespace alls;

namespace {

constexpr char kMeshMagic[8] = {'A', 'L', 'L', 'S', 'M', 'S', 'H', '1'};

// File header of a binary mesh (little-endian, 96 bytes).
struct MeshFileHeader {
    char magic[8];
    int32_t dim, nodes_per_elem, uniform_type, flags;   // flags: 1 elem_type, 2 E_per_elem
    int64_t n_nodes, n_elems;
    double E, nu, thickness;
    int64_t reserved[4];
};
static_assert(sizeof(MeshFileHeader) == 96, "mesh file header layout");

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

alls_status alls_csr_write_mtx(alls_ctx* ctx, const alls_csr* csr, const char* path) {
    if (!ctx || !csr || !path || !csr->row_ptr) return ALLS_E_INVALID;
    const int64_t n = csr->n_rows, nnz = csr->nnz;
    std::vector<int64_t> rp((size_t)n + 1);
    std::vector<int32_t> ci((size_t)std::max<int64_t>(nnz, 1));
    std::vector<double> v((size_t)std::max<int64_t>(nnz, 1));
    alls_status s = alls_csr_download(ctx, csr, rp.data(), ci.data(), v.data());
    if (s != ALLS_OK) return s;
    File out(path, "w");
    if (!out.f) {
        ctx->err = std::string("cannot open ") + path;
        return ALLS_E_INVALID;
    }
    std::fprintf(out.f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(out.f, "%% global stiffness matrix (A-LLS assembly, %s values)\n", csr->val_is_f32 ? "FP32" : "FP64");
    std::fprintf(out.f, "%lld %lld %lld\n", (long long)n, (long long)n, (long long)nnz);
    char line[64];
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            // %.17g round-trips every double exactly.
            const int len = std::snprintf(line, sizeof line, "%lld %d %.17g\n", (long long)(r + 1), ci[k] + 1, v[k]);
            if (std::fwrite(line, 1, (size_t)len, out.f) != (size_t)len) {
                ctx->err = std::string("write failed: ") + path;
                return ALLS_E_INVALID;
            }
        }
    return ALLS_OK;
}

alls_status alls_mesh_save(const alls_mesh* m, const char* path) {
    if (!m || !path || m->mem != ALLS_MEM_HOST) return ALLS_E_INVALID;
    if (m->n_nodes < 0 || m->n_elems < 0 || (m->n_nodes && !m->coords) || (m->n_elems && !m->conn))
        return ALLS_E_INVALID;
    MeshFileHeader h{};
    std::memcpy(h.magic, kMeshMagic, sizeof h.magic);
    h.dim = m->dim;
    h.nodes_per_elem = m->nodes_per_elem;
    h.uniform_type = m->uniform_type;
    h.flags = (m->elem_type ? 1 : 0) | (m->E_per_elem ? 2 : 0);
    h.n_nodes = m->n_nodes;
    h.n_elems = m->n_elems;
    h.E = m->E;
    h.nu = m->nu;
    h.thickness = m->thickness;
    File out(path, "wb");
    if (!out.f) return ALLS_E_INVALID;
    const size_t cb = (size_t)m->n_nodes * m->dim, kb = (size_t)m->n_elems * m->nodes_per_elem;
    bool ok = std::fwrite(&h, sizeof h, 1, out.f) == 1;
    ok = ok && std::fwrite(m->coords, sizeof(double), cb, out.f) == cb;
    ok = ok && std::fwrite(m->conn, sizeof(int32_t), kb, out.f) == kb;
    if (m->elem_type) ok = ok && std::fwrite(m->elem_type, 1, (size_t)m->n_elems, out.f) == (size_t)m->n_elems;
    if (m->E_per_elem)
        ok = ok && std::fwrite(m->E_per_elem, sizeof(double), (size_t)m->n_elems, out.f) == (size_t)m->n_elems;
    return ok ? ALLS_OK : ALLS_E_INVALID;
}

alls_status alls_mesh_load(const char* path, alls_host_mesh* out) {
    if (!path || !out) return ALLS_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    File in(path, "rb");
    if (!in.f) return ALLS_E_INVALID;
    MeshFileHeader h{};
    if (std::fread(&h, sizeof h, 1, in.f) != 1 || std::memcmp(h.magic, kMeshMagic, sizeof h.magic) != 0)
        return ALLS_E_INVALID;
    if ((h.dim != 2 && h.dim != 3) || h.nodes_per_elem < 3 || h.nodes_per_elem > 20 || h.n_nodes < 0 ||
        h.n_elems < 0 || h.n_nodes >= (int64_t)INT32_MAX || h.n_elems >= (int64_t)INT32_MAX)
        return ALLS_E_INVALID;
    const size_t cb = (size_t)h.n_nodes * h.dim, kb = (size_t)h.n_elems * h.nodes_per_elem;
    double* coords = static_cast<double*>(std::malloc(std::max<size_t>(cb, 1) * sizeof(double)));
    int32_t* conn = static_cast<int32_t*>(std::malloc(std::max<size_t>(kb, 1) * sizeof(int32_t)));
    uint8_t* types = (h.flags & 1) ? static_cast<uint8_t*>(std::malloc(std::max<int64_t>(h.n_elems, 1))) : nullptr;
    double* E = (h.flags & 2) ? static_cast<double*>(std::malloc(std::max<int64_t>(h.n_elems, 1) * sizeof(double)))
                              : nullptr;
    alls_mesh& m = out->mesh;
    m.coords = coords;
    m.conn = conn;
    m.elem_type = types;
    m.E_per_elem = E;
    if (!coords || !conn || ((h.flags & 1) && !types) || ((h.flags & 2) && !E)) {
        alls_host_mesh_free(out);
        return ALLS_E_OOM;
    }
    bool ok = std::fread(coords, sizeof(double), cb, in.f) == cb;
    ok = ok && std::fread(conn, sizeof(int32_t), kb, in.f) == kb;
    if (types) ok = ok && std::fread(types, 1, (size_t)h.n_elems, in.f) == (size_t)h.n_elems;
    if (E) ok = ok && std::fread(E, sizeof(double), (size_t)h.n_elems, in.f) == (size_t)h.n_elems;
    if (!ok) {
        alls_host_mesh_free(out);
        return ALLS_E_INVALID;
    }
    m.dim = h.dim;
    m.nodes_per_elem = h.nodes_per_elem;
    m.uniform_type = h.uniform_type;
    m.mem = ALLS_MEM_HOST;
    m.n_nodes = h.n_nodes;
    m.n_elems = h.n_elems;
    m.E = h.E;
    m.nu = h.nu;
    m.thickness = h.thickness;
    return ALLS_OK;
}

}  // extern "C"
