// cached_hex20.hip -- value-only Hex20 re-assembly on a cached pattern
// (gfx950), without atomics, keys or sorting.
//
// A repeat request integrates every element (k_hex20 with EMIT 3) and stores
// its 400 node-pair blocks in element order, blocks (I, J) of element e at
// stage[e*400 + I*20 + J]: contiguous, fully coalesced stores.  Then
// k_pair_gather assembles each node row a from its incident elements: the
// blocks (I_a, 0..19) of an element are contiguous, so every element
// contributes one 20-block segment (1.4 KB in FP64), and the cached rank map
// gives each block's column position.  The elements are taken in ascending
// id (the incidence list is sorted), three at a time for the loads and one
// after the other for the sums, so every column is summed in the serial
// order of Acc:336, exactly as the plain path sums it.
#include <algorithm>

#include "alls_internal.h"

namespace {

// Wave per node row; acc (LDS) holds the row's u column blocks.  A row of
// at most 3 * KS incident elements (every row of a conforming hexahedral
// mesh has at most 8) is taken in three dependent rounds of loads: the
// element ids (lane l: the row's l-th element), then their node lists
// (register slot k, lane (sub, J): node J of element 3k + sub), then each
// block's rank and values; the blocks are then added one element after the
// other.  Longer rows take three elements per round.
template <typename T, typename RP, int KS>
__global__ __launch_bounds__(64) void k_pair_gather(int64_t n_nodes, const int32_t* conn, const RP* rp,
                                                    const int32_t* rank_map, const int64_t* inc_off,
                                                    const int32_t* inc_list, const T* stage, double* val) {
    extern __shared__ double acc[];   // [u][9]
    const int lane = threadIdx.x;
    const int sub = lane / 20, J = lane - 20 * sub;   // element of the slot (0..2; 3 = idle), column node
    for (int64_t a = xcd_block(blockIdx.x, gridDim.x); a < n_nodes; a += gridDim.x) {
        const int64_t r0 = (int64_t)rp[3 * a], r1 = (int64_t)rp[3 * a + 1], r2 = (int64_t)rp[3 * a + 2];
        const int64_t u = (r1 - r0) / 3;
        const int64_t k0 = inc_off[a], k1 = inc_off[a + 1];
        const int64_t ne = k1 - k0;
        for (int64_t i = lane; i < u * 9; i += 64) acc[i] = 0.0;
        __syncthreads();
        if (ne <= 3 * KS) {
            const int32_t el = lane < ne ? inc_list[k0 + lane] : 0;
            int64_t e[KS];
            int32_t cj[KS];
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                const int idx = 3 * k + sub;
                e[k] = (int64_t)__shfl(el, idx < 64 ? idx : 0, 64);
                cj[k] = (sub < 3 && idx < ne) ? conn[e[k] * 20 + J] : -1;
            }
            int32_t r[KS];
            double v[KS][9];
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                // Local index of node a in element 3k + sub.
                const uint64_t hit = __ballot(cj[k] == a);
                const uint32_t mine = (uint32_t)(hit >> (20 * (sub < 3 ? sub : 0))) & 0xfffffu;
                const int I = (sub < 3 && mine) ? __ffs(mine) - 1 : -1;
                r[k] = -1;
#pragma unroll
                for (int q = 0; q < 9; ++q) v[k][q] = 0.0;
                if (I >= 0) {
                    const int64_t b = e[k] * 400 + I * 20 + J;
                    r[k] = rank_map[b];
                    const T* src = stage + b * 9;
#pragma unroll
                    for (int q = 0; q < 9; ++q) v[k][q] = (double)src[q];
                }
            }
            // In element order: one element's 20 blocks (distinct columns) per step.
#pragma unroll
            for (int k = 0; k < KS; ++k)
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (3 * k + s >= ne) break;
                    if (sub == s && r[k] >= 0) {
#pragma unroll
                        for (int q = 0; q < 9; ++q) acc[r[k] * 9 + q] += v[k][q];
                    }
                    __syncthreads();
                }
        } else {
            for (int64_t k = k0; k < k1; k += 3) {
                const int64_t kk = k + sub;
                const bool active = sub < 3 && kk < k1;
                int64_t e = 0;
                int32_t cj = -1;
                if (active) {
                    e = inc_list[kk];
                    cj = conn[e * 20 + J];
                }
                // Local index of node a in each of the three elements.
                const uint64_t hit = __ballot(active && cj == a);
                const uint32_t mine = (uint32_t)(hit >> (20 * (sub < 3 ? sub : 0))) & 0xfffffu;
                const int I = mine ? __ffs(mine) - 1 : -1;
                int32_t r = -1;
                double v[9];
                if (active && I >= 0) {
                    const int64_t b = e * 400 + I * 20 + J;
                    r = rank_map[b];
                    const T* src = stage + b * 9;
#pragma unroll
                    for (int q = 0; q < 9; ++q) v[q] = (double)src[q];
                }
                // In element order: the three elements add one after the other.
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (sub == s && r >= 0) {
#pragma unroll
                        for (int q = 0; q < 9; ++q) acc[r * 9 + q] += v[q];
                    }
                    __syncthreads();
                }
            }
        }
        const int64_t rows[3] = {r0, r1, r2};
        for (int64_t c = lane; c < u; c += 64) {
#pragma unroll
            for (int dr = 0; dr < 3; ++dr)
#pragma unroll
                for (int dc = 0; dc < 3; ++dc) val[rows[dr] + 3 * c + dc] = acc[c * 9 + dr * 3 + dc];
        }
        __syncthreads();
    }
}

// Largest number of column nodes in one node row: u = (rp[3a+1] - rp[3a]) / 3.
template <typename RP>
__global__ __launch_bounds__(256) void k_row_umax(int64_t n_nodes, const RP* rp, int32_t* out) {
    int32_t m = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_nodes; a += stride)
        m = max(m, (int32_t)(((int64_t)rp[3 * a + 1] - (int64_t)rp[3 * a]) / 3));
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

}  // namespace

namespace alls {

int32_t pair_gather_max_cols() { return 900; }   // 900 x 72 B of LDS

void launch_row_umax(int64_t n_nodes, const void* rp, int i64, int32_t* out, hipStream_t s) {
    if (n_nodes <= 0) return;
    int64_t g = (n_nodes + 255) / 256;
    if (g > 4096) g = 4096;
    if (i64)
        hipLaunchKernelGGL(k_row_umax<int64_t>, dim3((unsigned)g), dim3(256), 0, s, n_nodes,
                           static_cast<const int64_t*>(rp), out);
    else
        hipLaunchKernelGGL(k_row_umax<int32_t>, dim3((unsigned)g), dim3(256), 0, s, n_nodes,
                           static_cast<const int32_t*>(rp), out);
}

void launch_pair_gather(int64_t n_nodes, const int32_t* conn, const void* rp, int i64, const int32_t* rank_map,
                        const int64_t* inc_off, const int32_t* inc_list, const void* stage, int f32, int32_t umax,
                        double* val, hipStream_t s) {
    if (n_nodes <= 0) return;
    int64_t blocks = n_nodes;
    if (blocks > 262144) blocks = 262144;
    const size_t lds = (size_t)std::max(umax, 1) * 9 * sizeof(double);
#define ALLS_PG(T, RP)                                                                                      \
    hipLaunchKernelGGL((k_pair_gather<T, RP, 3>), dim3((unsigned)blocks), dim3(64), lds, s, n_nodes, conn,     \
                       static_cast<const RP*>(rp), rank_map, inc_off, inc_list, static_cast<const T*>(stage), val)
    if (f32) {
        if (i64) ALLS_PG(float, int64_t);
        else ALLS_PG(float, int32_t);
    } else {
        if (i64) ALLS_PG(double, int64_t);
        else ALLS_PG(double, int32_t);
    }
#undef ALLS_PG
}

}  // namespace alls
