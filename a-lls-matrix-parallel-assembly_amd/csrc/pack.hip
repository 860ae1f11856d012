// pack.hip -- batch packing (gfx950): n device-resident meshes of one
// dimension become one node/element space, mesh k's node ids shifted by the
// nodes of meshes 0..k-1, so that one A-LLS pipeline assembles all of them
// into a block-diagonal CSR.  Also the read-back of the batch's row_ptr at
// the mesh boundaries (nonzero offsets).
#include <algorithm>

#include "alls_internal.h"

namespace {

// Mesh holding global index i: the last k with off(k) <= i (offsets ascending).
template <bool NODES>
__device__ __forceinline__ int find_mesh(const alls::PackMesh* pm, int nm, int64_t i) {
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t o = NODES ? pm[mid].node_off : pm[mid].elem_off;
        if (o <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_pack_nodes(const alls::PackMesh* pm, int nm, int dim, int64_t n_nodes,
                                                    double* coords) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_nodes; i += stride) {
        const alls::PackMesh& m = pm[find_mesh<true>(pm, nm, i)];
        const int64_t li = i - m.node_off;
        for (int d = 0; d < dim; ++d) coords[i * dim + d] = m.coords[li * dim + d];
    }
}

// Element e: connectivity shifted by the mesh's node offset (-1 padding kept,
// slots beyond the source stride become -1), element type and Young's
// modulus when the packed mesh carries them.
__global__ __launch_bounds__(256) void k_pack_elems(const alls::PackMesh* pm, int nm, int npe, int64_t n_elems,
                                                    int32_t* conn, uint8_t* types, double* E) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_elems; e += stride) {
        const alls::PackMesh& m = pm[find_mesh<false>(pm, nm, e)];
        const int64_t le = e - m.elem_off;
        for (int j = 0; j < npe; ++j) {
            int32_t v = -1;
            if (j < m.npe) {
                const int32_t c = m.conn[le * m.npe + j];
                v = c >= 0 ? c + (int32_t)m.node_off : -1;
            }
            conn[e * npe + j] = v;
        }
        if (types) types[e] = m.types ? m.types[le] : (uint8_t)m.type;
        if (E) E[e] = m.E_elem ? m.E_elem[le] : m.E;
    }
}

template <typename RP>
__global__ void k_gather_rp(const RP* rp, const int64_t* idx, int n, int64_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int64_t)rp[idx[i]];
}

// Request k of the block-diagonal batch result: global rows
// [row_off, row_off+n_rows) and nonzeros [nnz_off, nnz_off+nnz) go to its
// own CSR with local numbering (row_ptr - nnz_off, col_ind - row_off).
template <bool BY_ROW>
__device__ __forceinline__ int find_part(const alls::SplitPart* pt, int n, int64_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((BY_ROW ? pt[mid].row_off : pt[mid].nnz_off) <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename RP>
__global__ __launch_bounds__(256) void k_split(const alls::SplitPart* pt, int n, int64_t total_rows,
                                               int64_t total_nnz, const RP* rp, const int32_t* ci, const double* v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t j = t0; j < total_nnz; j += stride) {
        const alls::SplitPart& p = pt[find_part<false>(pt, n, j)];
        const int64_t lj = j - p.nnz_off;
        p.col_ind[lj] = ci[j] - (int32_t)p.row_off;
        p.val[lj] = v[j];
    }
    for (int64_t i = t0; i < total_rows; i += stride) {
        const alls::SplitPart& p = pt[find_part<true>(pt, n, i)];
        p.row_ptr[i - p.row_off] = (int32_t)((int64_t)rp[i] - p.nnz_off);
    }
    for (int64_t k = t0; k < n; k += stride) pt[k].row_ptr[pt[k].n_rows] = (int32_t)pt[k].nnz;
}

unsigned grid_cap(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

namespace alls {

void launch_pack(const PackMesh* pm, int nm, int dim, int npe, int64_t n_nodes, int64_t n_elems, double* coords,
                 int32_t* conn, uint8_t* types, double* E, hipStream_t s) {
    if (n_nodes > 0)
        hipLaunchKernelGGL(k_pack_nodes, dim3(grid_cap(n_nodes)), dim3(256), 0, s, pm, nm, dim, n_nodes, coords);
    if (n_elems > 0)
        hipLaunchKernelGGL(k_pack_elems, dim3(grid_cap(n_elems)), dim3(256), 0, s, pm, nm, npe, n_elems, conn, types,
                           E);
}

void launch_split(const SplitPart* parts, int n, int64_t total_rows, int64_t total_nnz, const void* row_ptr,
                  int i64, const int32_t* col_ind, const double* val, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = grid_cap(std::max<int64_t>(std::max(total_nnz, total_rows), n));
    if (i64)
        hipLaunchKernelGGL(k_split<int64_t>, dim3(g), dim3(256), 0, s, parts, n, total_rows, total_nnz,
                           static_cast<const int64_t*>(row_ptr), col_ind, val);
    else
        hipLaunchKernelGGL(k_split<int32_t>, dim3(g), dim3(256), 0, s, parts, n, total_rows, total_nnz,
                           static_cast<const int32_t*>(row_ptr), col_ind, val);
}

void launch_gather_rp(const void* row_ptr, int i64, const int64_t* idx, int n, int64_t* out, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (i64)
        hipLaunchKernelGGL(k_gather_rp<int64_t>, dim3(g), dim3(256), 0, s, static_cast<const int64_t*>(row_ptr), idx,
                           n, out);
    else
        hipLaunchKernelGGL(k_gather_rp<int32_t>, dim3(g), dim3(256), 0, s, static_cast<const int32_t*>(row_ptr), idx,
                           n, out);
}

}  // namespace alls
