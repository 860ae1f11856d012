// meshgen.cpp -- host generators of the synthetic benchmark meshes.
//
// Structured grids as in the reference (generateMesh Acc:575-603,
// MeshGenerator T3:366-456 / Mix:393-486, generateMeshHex20 Hex:324-362),
// with: interior-node jitter, random split diagonals, random node-id
// permutation, lognormal per-element E, and compact Hex20 numbering (no
// orphan face/body-centre nodes, unlike the reference's full lattice).
// All randomness comes from a counter-based SplitMix64 hash of (seed,
// stream, index), so a mesh is reproducible from its spec alone.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "alls.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Uniform in [0,1) from (seed, stream, index).
inline double u01(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull) ^ mix64(i));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
// Standard normal by Box-Muller from two uniforms of the same index.
inline double normal(uint64_t seed, uint64_t stream, uint64_t i) {
    double a = 1.0 - u01(seed, stream, 2 * i);        // (0,1]
    double b = u01(seed, stream, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

enum : uint64_t { S_JITTER = 1, S_DIAG = 2, S_SPLIT = 3, S_PERM = 4, S_E = 5 };

template <class T> T* dup(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T) + 16));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

// Random permutation of [0,n) (Fisher-Yates on the counter RNG).
std::vector<int32_t> permutation(int64_t n, uint64_t seed) {
    std::vector<int32_t> p((size_t)n);
    for (int64_t i = 0; i < n; ++i) p[i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; --i) {
        int64_t j = (int64_t)(u01(seed, S_PERM, (uint64_t)i) * (double)(i + 1));
        if (j > i) j = i;
        std::swap(p[i], p[j]);
    }
    return p;
}

}  // namespace

extern "C" {

alls_status alls_gen_mesh(const alls_gen_spec* sp, alls_host_mesh* out) {
    if (!sp || !out) return ALLS_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    const int nx = sp->nx, ny = sp->ny, nz = sp->nz;
    if (nx <= 0 || ny <= 0 || (sp->kind == ALLS_GEN_HEX20_BRICK && nz <= 0)) return ALLS_E_INVALID;
    const uint64_t seed = sp->seed;
    std::vector<double> coords;
    std::vector<int32_t> conn, color;
    std::vector<uint8_t> types;
    int64_t n_nodes = 0;
    int dim = 2, npe = 4;
    int ncolors = 4;
    int uniform = ALLS_Q4;

    if (sp->kind == ALLS_GEN_HEX20_BRICK) {
        dim = 3;
        npe = 20;
        uniform = ALLS_HEX20;
        ncolors = 8;
        // Half-step lattice (2nx+1)(2ny+1)(2nz+1); keep points with at most one
        // odd coordinate (corners and mid-edge nodes), numbered in (k,j,i) order.
        const int64_t LX = 2 * nx + 1, LY = 2 * ny + 1, LZ = 2 * nz + 1;
        std::vector<int32_t> id((size_t)(LX * LY * LZ), -1);
        for (int64_t k = 0; k < LZ; ++k)
            for (int64_t j = 0; j < LY; ++j)
                for (int64_t i = 0; i < LX; ++i) {
                    if ((i & 1) + (j & 1) + (k & 1) > 1) continue;
                    const int64_t lid = (k * LY + j) * LX + i;
                    id[lid] = (int32_t)n_nodes++;
                    double p[3] = {0.5 * i / nx, 0.5 * j / ny, 0.5 * k / nz};
                    const bool interior = i > 0 && i < LX - 1 && j > 0 && j < LY - 1 && k > 0 && k < LZ - 1;
                    if (interior && sp->jitter > 0) {
                        const double h[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
                        for (int d = 0; d < 3; ++d)
                            p[d] += (2.0 * u01(seed, S_JITTER, (uint64_t)lid * 3 + d) - 1.0) * sp->jitter * h[d];
                    }
                    coords.push_back(p[0]);
                    coords.push_back(p[1]);
                    coords.push_back(p[2]);
                }
        auto L = [&](int64_t i, int64_t j, int64_t k) { return id[(size_t)((k * LY + j) * LX + i)]; };
        // Reference node order (Hex:343-348): corners bottom/top, mid-edges
        // bottom/top, then vertical mid-edges.
        static const int off[20][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2},
                                       {0, 2, 2}, {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2},
                                       {1, 2, 2}, {0, 1, 2}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}};
        conn.reserve((size_t)nx * ny * nz * 20);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    for (int n = 0; n < 20; ++n) conn.push_back(L(2 * i + off[n][0], 2 * j + off[n][1], 2 * k + off[n][2]));
                    color.push_back((i & 1) + 2 * (j & 1) + 4 * (k & 1));
                }
    } else {
        const int64_t NX = nx + 1;
        n_nodes = (int64_t)(nx + 1) * (ny + 1);
        coords.resize((size_t)n_nodes * 2);
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i) {
                const int64_t nid = j * NX + i;
                double x = (double)i / nx, y = (double)j / ny;
                if (i > 0 && i < nx && j > 0 && j < ny && sp->jitter > 0) {
                    x += (2.0 * u01(seed, S_JITTER, (uint64_t)nid * 2) - 1.0) * sp->jitter / nx;
                    y += (2.0 * u01(seed, S_JITTER, (uint64_t)nid * 2 + 1) - 1.0) * sp->jitter / ny;
                }
                coords[2 * nid] = x;
                coords[2 * nid + 1] = y;
            }
        const bool mixed = sp->kind == ALLS_GEN_MIXED_GRID;
        const bool t3 = sp->kind == ALLS_GEN_T3_GRID;
        if (sp->kind == ALLS_GEN_Q4_GRID) {
            uniform = ALLS_Q4;
            npe = 4;
        } else if (t3) {
            uniform = ALLS_T3;
            npe = 3;
            ncolors = 8;
        } else if (mixed) {
            npe = 4;
            ncolors = 8;
        } else {
            return ALLS_E_INVALID;
        }
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                // CCW corners as the reference (Acc:592-595).
                const int32_t n0 = (int32_t)(j * NX + i), n1 = n0 + 1;
                const int32_t n2 = (int32_t)((j + 1) * NX + i + 1), n3 = n2 - 1;
                const int cc = (int)((i & 1) + 2 * (j & 1));
                const uint64_t cell = (uint64_t)(j * nx + i);
                bool split = t3;
                if (mixed) split = u01(seed, S_SPLIT, cell) < sp->split_fraction;
                if (!split) {
                    conn.insert(conn.end(), {n0, n1, n2, n3});
                    if (mixed) types.push_back(ALLS_Q4);
                    color.push_back(mixed ? 2 * cc : cc);
                    continue;
                }
                const bool diag13 = sp->random_diagonal && u01(seed, S_DIAG, cell) < 0.5;
                int32_t a[3], b[3];
                if (!diag13) {   // along n0-n2, as T3:387-390
                    a[0] = n0; a[1] = n1; a[2] = n2;
                    b[0] = n0; b[1] = n2; b[2] = n3;
                } else {         // along n1-n3
                    a[0] = n0; a[1] = n1; a[2] = n3;
                    b[0] = n1; b[1] = n2; b[2] = n3;
                }
                // The two triangles of a cell share an edge: colour 2*cc + t.
                int t = 0;
                for (const int32_t* tri : {a, b}) {
                    conn.insert(conn.end(), {tri[0], tri[1], tri[2]});
                    if (mixed) {
                        conn.push_back(-1);
                        types.push_back(ALLS_T3);
                    }
                    color.push_back(2 * cc + t++);
                }
            }
    }
    const int64_t n_elems = (int64_t)(conn.size() / npe);
    if (sp->permute_nodes) {
        std::vector<int32_t> perm = permutation(n_nodes, seed);
        std::vector<double> c2(coords.size());
        for (int64_t a = 0; a < n_nodes; ++a)
            for (int d = 0; d < dim; ++d) c2[(size_t)perm[a] * dim + d] = coords[(size_t)a * dim + d];
        coords.swap(c2);
        for (auto& c : conn)
            if (c >= 0) c = perm[c];
    }
    std::vector<double> E;
    if (sp->E_sigma > 0) {
        E.resize((size_t)n_elems);
        for (int64_t e = 0; e < n_elems; ++e) E[e] = std::exp(sp->E_sigma * normal(seed, S_E, (uint64_t)e));
    }
    alls_mesh& m = out->mesh;
    m.dim = dim;
    m.nodes_per_elem = npe;
    m.n_nodes = n_nodes;
    m.n_elems = n_elems;
    m.coords = dup(coords);
    m.conn = dup(conn);
    m.elem_type = types.empty() ? nullptr : dup(types);
    m.uniform_type = uniform;
    m.mem = ALLS_MEM_HOST;
    m.E = sp->E;
    m.E_per_elem = E.empty() ? nullptr : dup(E);
    m.nu = sp->nu;
    m.thickness = sp->thickness;
    out->structured_color = dup(color);
    out->n_structured_colors = ncolors;
    if (!m.coords || !m.conn || !out->structured_color) {
        alls_host_mesh_free(out);
        return ALLS_E_OOM;
    }
    return ALLS_OK;
}

void alls_host_mesh_free(alls_host_mesh* h) {
    if (!h) return;
    std::free(const_cast<double*>(h->mesh.coords));
    std::free(const_cast<int32_t*>(h->mesh.conn));
    std::free(const_cast<uint8_t*>(h->mesh.elem_type));
    std::free(const_cast<double*>(h->mesh.E_per_elem));
    std::free(h->structured_color);
    std::memset(h, 0, sizeof(*h));
}

}  // extern "C"
