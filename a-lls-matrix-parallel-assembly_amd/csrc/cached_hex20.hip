// cached_hex20.hip -- value-only Hex20 re-assembly on a cached pattern
// (gfx950), without atomics, keys or sorting.
//
// At pattern build time every contribution (element e, row node I, column
// node J) gets a fixed slot in a staging array grouped by node pair
// (a = conn[e][I], rank r of b = conn[e][J] in row a), elements inside a group
// in ascending id (the serial order of Acc:336):
//   k_pair_count : contributions per node pair (elements sharing a and b);
//   (scan)       : group offsets;
//   k_slot_map   : slot of each contribution.
// Both walk each node row's incident elements with one thread, which owns
// the row's node-pair counters, so they need no atomics.
// A repeat request then integrates each element (k_hex20 with EMIT 3 writes
// its 400 blocks to their slots) and k_pair_sum adds every group and writes
// the 3x3 block into the CSR rows.
#include "alls_internal.h"

namespace {

// Local index of node a in element e (a is one of its 20 nodes).
__device__ __forceinline__ int local_index(const int32_t* conn, int64_t e, int64_t a) {
    const int32_t* c = conn + e * 20;
    int I = 0;
#pragma unroll
    for (int i = 0; i < 20; ++i)
        if (c[i] == a) I = i;
    return I;
}

// Thread per node row a: the row owns its node pairs g0 .. g0+u-1, so the
// counters need no atomics.  For every element incident to a, every one of
// its 20 column nodes adds one contribution to pair (a, rank).
template <typename RP>
__global__ __launch_bounds__(256) void k_pair_count(int64_t n_nodes, const int32_t* conn, const RP* rp,
                                                    const int32_t* rank_map, const int64_t* inc_off,
                                                    const int32_t* inc_list, int32_t* cnt) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_nodes) return;
    const int64_t r0 = (int64_t)rp[3 * a];
    const int64_t u = ((int64_t)rp[3 * a + 1] - r0) / 3;
    const int64_t g0 = r0 / 9;
    for (int64_t r = 0; r < u; ++r) cnt[g0 + r] = 0;
    for (int64_t k = inc_off[a]; k < inc_off[a + 1]; ++k) {
        const int64_t e = inc_list[k];
        const int32_t* mp = rank_map + e * 400 + local_index(conn, e, a) * 20;
        for (int J = 0; J < 20; ++J) cnt[g0 + mp[J]] += 1;
    }
}

// Thread per node row a, elements in ascending id (the incidence list is
// sorted): contribution (e, I, J) takes the next free slot of its pair's
// group, so groups come out in the serial order of Acc:336.  `fill` holds
// the row's running counts (zeroed here).
template <typename RP>
__global__ __launch_bounds__(256) void k_slot_map(int64_t n_nodes, const int32_t* conn, const RP* rp,
                                                  const int32_t* rank_map, const int64_t* inc_off,
                                                  const int32_t* inc_list, const int64_t* goff, int32_t* fill,
                                                  int64_t* slot) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_nodes) return;
    const int64_t r0 = (int64_t)rp[3 * a];
    const int64_t u = ((int64_t)rp[3 * a + 1] - r0) / 3;
    const int64_t g0 = r0 / 9;
    for (int64_t r = 0; r < u; ++r) fill[g0 + r] = 0;
    for (int64_t k = inc_off[a]; k < inc_off[a + 1]; ++k) {
        const int64_t e = inc_list[k];
        const int I = local_index(conn, e, a);
        const int32_t* mp = rank_map + e * 400 + I * 20;
        for (int J = 0; J < 20; ++J) {
            const int64_t g = g0 + mp[J];
            slot[e * 400 + I * 20 + J] = goff[g] + fill[g]++;
        }
    }
}

// Wave per node row: sums each node pair's contributions in slot order and
// writes the 3x3 block to DOF rows 3a..3a+2.
template <typename T, typename RP>
__global__ __launch_bounds__(256) void k_pair_sum(int64_t n_nodes, const RP* rp, const int64_t* goff, const T* stage,
                                                  double* val) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t a = wave; a < n_nodes; a += nw) {
        const int64_t r0 = (int64_t)rp[3 * a], r1 = (int64_t)rp[3 * a + 1], r2 = (int64_t)rp[3 * a + 2];
        const int64_t u = (r1 - r0) / 3;
        const int64_t g0 = r0 / 9;
        for (int64_t r = lane; r < u; r += 64) {
            double acc[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] = 0.0;
            for (int64_t s = goff[g0 + r]; s < goff[g0 + r + 1]; ++s) {
#pragma unroll
                for (int v = 0; v < 9; ++v) acc[v] += (double)stage[s * 9 + v];
            }
            const int64_t rows[3] = {r0, r1, r2};
#pragma unroll
            for (int dr = 0; dr < 3; ++dr)
#pragma unroll
                for (int dc = 0; dc < 3; ++dc) val[rows[dr] + 3 * r + dc] = acc[dr * 3 + dc];
        }
    }
}

unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

namespace alls {

void launch_pair_count(int64_t n_nodes, const int32_t* conn, const void* rp, int i64, const int32_t* rank_map,
                       const int64_t* inc_off, const int32_t* inc_list, int32_t* cnt, hipStream_t s) {
    if (n_nodes <= 0) return;
    if (i64)
        hipLaunchKernelGGL(k_pair_count<int64_t>, dim3(grid_for(n_nodes)), dim3(256), 0, s, n_nodes, conn,
                           static_cast<const int64_t*>(rp), rank_map, inc_off, inc_list, cnt);
    else
        hipLaunchKernelGGL(k_pair_count<int32_t>, dim3(grid_for(n_nodes)), dim3(256), 0, s, n_nodes, conn,
                           static_cast<const int32_t*>(rp), rank_map, inc_off, inc_list, cnt);
}

void launch_slot_map(int64_t n_nodes, const int32_t* conn, const void* rp, int i64, const int32_t* rank_map,
                     const int64_t* inc_off, const int32_t* inc_list, const int64_t* goff, int32_t* fill,
                     int64_t* slot, hipStream_t s) {
    if (n_nodes <= 0) return;
    if (i64)
        hipLaunchKernelGGL(k_slot_map<int64_t>, dim3(grid_for(n_nodes)), dim3(256), 0, s, n_nodes, conn,
                           static_cast<const int64_t*>(rp), rank_map, inc_off, inc_list, goff, fill, slot);
    else
        hipLaunchKernelGGL(k_slot_map<int32_t>, dim3(grid_for(n_nodes)), dim3(256), 0, s, n_nodes, conn,
                           static_cast<const int32_t*>(rp), rank_map, inc_off, inc_list, goff, fill, slot);
}

void launch_pair_sum(int64_t n_nodes, const void* rp, int i64, const int64_t* goff, const void* stage, int f32,
                     double* val, hipStream_t s) {
    if (n_nodes <= 0) return;
    int64_t blocks = (n_nodes + 3) / 4;
    if (blocks > 262144) blocks = 262144;
#define ALLS_PS(T, RP)                                                                                        \
    hipLaunchKernelGGL((k_pair_sum<T, RP>), dim3((unsigned)blocks), dim3(256), 0, s, n_nodes,                 \
                       static_cast<const RP*>(rp), goff, static_cast<const T*>(stage), val)
    if (f32) {
        if (i64) ALLS_PS(float, int64_t);
        else ALLS_PS(float, int32_t);
    } else {
        if (i64) ALLS_PS(double, int64_t);
        else ALLS_PS(double, int32_t);
    }
#undef ALLS_PS
}

}  // namespace alls
