// color.hip -- element colouring on the GPU and colour bucketing (gfx950).
//
// Replaces the host greedy colourers (Acc:648-747 degree-ordered,
// T3:484-523 and Hex:394-417 natural order), which are O(N^2) in memory
// (a std::vector<bool> of N per element).  Elements are adjacent when they
// share a node.  Jones-Plassmann-Luby rounds: an uncoloured element whose
// (hash priority, id) beats every uncoloured neighbour takes the smallest
// colour no coloured neighbour has.  Each round reads the previous round's
// colours and writes a new array, so the result does not depend on thread
// timing; two adjacent elements never colour in the same round.
#include <algorithm>

#include "alls_internal.h"

namespace {

__device__ __forceinline__ uint32_t prio(uint32_t e) {
    uint32_t h = e * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ int elem_nodes(const DevMesh& m, int64_t e) {
    const int kind = m.types ? (int)m.types[e] : m.uniform_type;
    return kind == 0 ? 4 : kind == 1 ? 3 : 20;
}

__global__ void k_inc_count(DevMesh m, int32_t* cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.n_elems) return;
    const int nn = elem_nodes(m, e);
    for (int i = 0; i < nn; ++i) atomicAdd(&cnt[m.conn[e * m.npe + i]], 1);
}

__global__ void k_inc_fill(DevMesh m, const int64_t* off, int32_t* cursor, int32_t* list) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.n_elems) return;
    const int nn = elem_nodes(m, e);
    for (int i = 0; i < nn; ++i) {
        const int32_t a = m.conn[e * m.npe + i];
        list[off[a] + atomicAdd(&cursor[a], 1)] = (int32_t)e;
    }
}

// Sorts each node's incident elements by id (insertion sort; lists are short)
// so that gathers visit elements in the serial assembly's order.
__global__ void k_inc_sort(int64_t n_nodes, const int64_t* off, int32_t* list) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_nodes) return;
    const int64_t lo = off[a], hi = off[a + 1];
    for (int64_t i = lo + 1; i < hi; ++i) {
        const int32_t v = list[i];
        int64_t j = i - 1;
        while (j >= lo && list[j] > v) {
            list[j + 1] = list[j];
            --j;
        }
        list[j + 1] = v;
    }
}

// Adds a wave's flags and largest colour into misc[0] / misc[1] with one
// atomic each from lane 0 (every lane of the wave calls it): per-element
// atomics on these two words serialised the rounds, all lanes of the grid
// queuing on one address.
__device__ __forceinline__ void wave_flush(bool flag, int32_t colour, int32_t* misc) {
    const uint64_t fm = __ballot(flag);
    int32_t mx = colour;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
    if ((threadIdx.x & 63) == 0) {
        if (fm) atomicAdd(&misc[0], (int32_t)__popcll(fm));
        if (mx >= 0) atomicMax(&misc[1], mx);
    }
}

// One JPL round.  misc[0] counts elements still uncoloured after the round,
// misc[1] is the largest colour used.
__global__ __launch_bounds__(256) void k_jpl_round(DevMesh m, const int64_t* off, const int32_t* list,
                                                   const int32_t* cin, int32_t* cout, int32_t* misc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool pending = false;     // still uncoloured after this round
    int32_t taken = -1;       // colour taken in this round
    if (e < m.n_elems) {
        const int32_t ce = cin[e];
        if (ce >= 0) {
            cout[e] = ce;
        } else {
            const uint32_t pe = prio((uint32_t)e);
            const int nn = elem_nodes(m, e);
            bool best = true;
            for (int i = 0; i < nn && best; ++i) {
                const int32_t a = m.conn[e * m.npe + i];
                for (int64_t k = off[a]; k < off[a + 1]; ++k) {
                    const int32_t f = list[k];
                    if (f == e || cin[f] >= 0) continue;
                    const uint32_t pf = prio((uint32_t)f);
                    if (pf > pe || (pf == pe && f > e)) {
                        best = false;
                        break;
                    }
                }
            }
            if (!best) {
                cout[e] = -1;
                pending = true;
            } else {
                // Smallest colour absent from the coloured neighbours, 64 colours at a time.
                for (int32_t base = 0; taken < 0; base += 64) {
                    uint64_t used = 0;
                    for (int i = 0; i < nn; ++i) {
                        const int32_t a = m.conn[e * m.npe + i];
                        for (int64_t k = off[a]; k < off[a + 1]; ++k) {
                            const int32_t c = cin[list[k]] - base;
                            if (c >= 0 && c < 64) used |= 1ull << c;
                        }
                    }
                    if (~used) taken = base + __ffsll((long long)~used) - 1;
                }
                cout[e] = taken;
            }
        }
    }
    wave_flush(pending, taken, misc);
}

// Counts colour conflicts (adjacent elements with equal colour) and negative
// colours into misc[0], the largest colour into misc[1].
__global__ __launch_bounds__(256) void k_conflicts(DevMesh m, const int64_t* off, const int32_t* list,
                                                   const int32_t* col, int32_t* misc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    int32_t ce = -1;
    if (e < m.n_elems) {
        ce = col[e];
        bad = ce < 0;
        const int nn = bad ? 0 : elem_nodes(m, e);
        for (int i = 0; i < nn && !bad; ++i) {
            const int32_t a = m.conn[e * m.npe + i];
            for (int64_t k = off[a]; k < off[a + 1]; ++k) {
                const int32_t f = list[k];
                if (f != e && col[f] == ce) {
                    bad = true;
                    break;
                }
            }
        }
    }
    wave_flush(bad, ce, misc);
}

// Colour histogram of nc <= COLOR_LDS colours: counts per block in LDS,
// then one global add per colour per block (a global add per element queued
// the whole grid on a dozen words: 12.9 ms for 4M elements).
constexpr int COLOR_LDS = 1024;
constexpr int COLOR_EPT = 4;   // elements per thread and pass of the bucket kernel

__global__ __launch_bounds__(256) void k_color_count(const int32_t* col, int64_t n, int nc, int32_t* cnt) {
    __shared__ int32_t h[COLOR_LDS];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) atomicAdd(&h[col[e]], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += blockDim.x)
        if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// Buckets elements by colour.  Per pass a block takes COLOR_EPT x 256
// elements: ranks inside the block come from LDS atomics, then one global
// add per colour reserves the block's range.  The order inside a colour is
// arbitrary: an entry receives at most one contribution per colour, so
// coloured sums do not depend on it.
__global__ __launch_bounds__(256) void k_color_bucket(const int32_t* col, int64_t n, int nc, const int64_t* coff,
                                                      int32_t* cursor, int32_t* elist) {
    __shared__ int32_t h[COLOR_LDS];
    const int64_t per_pass = (int64_t)COLOR_EPT * blockDim.x;
    for (int64_t t0 = (int64_t)blockIdx.x * per_pass; t0 < n; t0 += (int64_t)gridDim.x * per_pass) {
        for (int i = threadIdx.x; i < nc; i += blockDim.x) h[i] = 0;
        __syncthreads();
        int32_t c[COLOR_EPT], r[COLOR_EPT];
#pragma unroll
        for (int k = 0; k < COLOR_EPT; ++k) {
            const int64_t e = t0 + k * blockDim.x + threadIdx.x;
            c[k] = e < n ? col[e] : -1;
            r[k] = c[k] >= 0 ? atomicAdd(&h[c[k]], 1) : 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nc; i += blockDim.x)
            if (h[i]) h[i] = atomicAdd(&cursor[i], h[i]);   // the block's first slot of colour i
        __syncthreads();
#pragma unroll
        for (int k = 0; k < COLOR_EPT; ++k) {
            const int64_t e = t0 + k * blockDim.x + threadIdx.x;
            if (c[k] >= 0) elist[coff[c[k]] + h[c[k]] + r[k]] = (int32_t)e;
        }
        __syncthreads();
    }
}

// Same for more colours than COLOR_LDS (one global add per element).
__global__ void k_color_count_wide(const int32_t* col, int64_t n, int32_t* cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) atomicAdd(&cnt[col[e]], 1);
}

__global__ void k_color_bucket_wide(const int32_t* col, int64_t n, const int64_t* coff, int32_t* cursor,
                                    int32_t* elist) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int32_t c = col[e];
    elist[coff[c] + atomicAdd(&cursor[c], 1)] = (int32_t)e;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

namespace alls {

void launch_incidence(const DevMesh& m, int32_t* cnt, int64_t* off, int32_t* cursor, int32_t* list, int64_t* tiles,
                      hipStream_t s) {
    launch_fill_i32(cnt, m.n_nodes, 0, s);
    launch_fill_i32(cursor, m.n_nodes, 0, s);
    if (m.n_elems > 0) hipLaunchKernelGGL(k_inc_count, dim3(grid_for(m.n_elems)), dim3(256), 0, s, m, cnt);
    launch_scan_total(cnt, m.n_nodes, tiles, scan_tile_count(m.n_nodes), s);
    launch_scan_apply(cnt, m.n_nodes, tiles, 1, off, 1, s);
    if (m.n_elems > 0) hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(m.n_elems)), dim3(256), 0, s, m, off, cursor, list);
    if (m.n_nodes > 0) hipLaunchKernelGGL(k_inc_sort, dim3(grid_for(m.n_nodes)), dim3(256), 0, s, m.n_nodes, off, list);
}

void launch_jpl_round(const DevMesh& m, const int64_t* off, const int32_t* list, const int32_t* cin, int32_t* cout,
                      int32_t* misc, hipStream_t s) {
    hipLaunchKernelGGL(k_jpl_round, dim3(grid_for(m.n_elems)), dim3(256), 0, s, m, off, list, cin, cout, misc);
}

void launch_conflicts(const DevMesh& m, const int64_t* off, const int32_t* list, const int32_t* col, int32_t* misc,
                      hipStream_t s) {
    if (m.n_elems > 0) hipLaunchKernelGGL(k_conflicts, dim3(grid_for(m.n_elems)), dim3(256), 0, s, m, off, list, col, misc);
}

void launch_color_count(const int32_t* col, int64_t n, int nc, int32_t* cnt, hipStream_t s) {
    if (n <= 0) return;
    if (nc > COLOR_LDS) {
        hipLaunchKernelGGL(k_color_count_wide, dim3(grid_for(n)), dim3(256), 0, s, col, n, cnt);
        return;
    }
    const unsigned g = std::min<unsigned>(grid_for(n), 2048);
    hipLaunchKernelGGL(k_color_count, dim3(g), dim3(256), 0, s, col, n, nc, cnt);
}

void launch_color_bucket(const int32_t* col, int64_t n, int nc, const int64_t* coff, int32_t* cursor,
                         int32_t* elist, hipStream_t s) {
    if (n <= 0) return;
    if (nc > COLOR_LDS) {
        hipLaunchKernelGGL(k_color_bucket_wide, dim3(grid_for(n)), dim3(256), 0, s, col, n, coff, cursor, elist);
        return;
    }
    const unsigned g = std::min<unsigned>(grid_for((n + COLOR_EPT - 1) / COLOR_EPT), 2048);
    hipLaunchKernelGGL(k_color_bucket, dim3(g), dim3(256), 0, s, col, n, nc, coff, cursor, elist);
}

}  // namespace alls
