// consolidate.hip -- A-LLS lists -> CSR on the device (gfx950).
//
// The reference builds row_ptr/col_ind on the host with std::set
// (Acc:605-645) and its convertALLStoCSRKernel (Acc:471-501) walks each row's
// list with one thread, binary-searching every node's column in the host
// pattern.  Here the pattern itself is produced on the device:
//   count : per node row, sort its (column, element) keys and count unique
//           columns u[a];
//   scan  : u -> row_ptr (tile sums, one-block scan of the tile sums, apply);
//   write : sort again, merge equal columns (sum in element-id order) and
//           write col_ind / val of the dpn DOF rows of the node.
// Rows of <= G entries use G lanes and a register bitonic sort through
// cross-lane shuffles.  Longer rows (Hex20, or unstructured outliers) use a
// wave per row: an LDS hash set for the count, a sort staged in LDS for the
// write.  Sorting by (column, element id)
// makes the summation order, and so the values, independent of the order in
// which atomics appended the entries.
#include "alls_internal.h"

namespace {

constexpr uint64_t KEY_INF = ~0ull;

__device__ __forceinline__ uint64_t pack_key(int2 kv) {
    return ((uint64_t)(uint32_t)kv.x << 32) | (uint32_t)kv.y;
}
__device__ __forceinline__ int32_t key_col(uint64_t k) { return (int32_t)(k >> 32); }

__device__ __forceinline__ int64_t entry_index(const ListPool& p, int64_t a, const int32_t* h, int32_t s) {
    // Callers clamp s below ALLS_DIR * C.  A chunk that was never published
    // (only after a directory overflow, which the host then retries) maps to
    // the home chunk so that every access stays in bounds.
    const int32_t k = s / p.C;
    const int32_t d = k == 0 ? 0 : h[k];
    const int64_t ch = d > 0 ? p.n_nodes + (int64_t)(d - 1) : a;
    return ch * p.C + (s % p.C);
}

// ------------------------------------------------ register (small) rows ---
template <int G, int B, typename VT, bool WRITE, typename RP>
__global__ __launch_bounds__(256) void k_rows_small(ListPool p, int32_t* u, int32_t* big_list,
                                                    const RP* row_ptr, int32_t* col_ind, double* val,
                                                    int dpn) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a = gid / G;
    const int lane = (int)(gid % G);
    if (a >= p.n_nodes) return;
    const int32_t* h = p.hdr + a * ALLS_DIR;
    const int32_t n = min(h[0], ALLS_DIR * p.C);
    if (n > G) {
        if (!WRITE && lane == 0) big_list[atomicAdd(&p.misc[MISC_BIG_COUNT], 1)] = (int32_t)a;
        return;
    }
    uint64_t key = KEY_INF;
    int32_t slot = lane;
    if (lane < n) key = pack_key(p.keys[entry_index(p, a, h, lane)]);
    // Bitonic sort of G keys across the group's lanes.
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t ok = __shfl_xor(key, j, G);
            const int32_t os = __shfl_xor(slot, j, G);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            const bool take = (lower == up) ? (ok < key) : (ok > key);
            if (take) {
                key = ok;
                slot = os;
            }
        }
    }
    const int32_t col = key_col(key);
    const bool valid = lane < n;
    const int32_t pcol = key_col(__shfl_up(key, 1, G));
    const bool head = valid && (lane == 0 || pcol != col);
    const uint64_t heads = __ballot(head);
    const int wl = threadIdx.x & 63;
    const int gbase = wl - lane;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gbase);
    const uint64_t gheads = heads & gmask;
    if (!WRITE) {
        if (lane == 0) u[a] = __popcll(gheads);
        return;
    }
    double v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) v[b] = 0.0;
    if (valid) {
        const VT* src = static_cast<const VT*>(p.vals) + entry_index(p, a, h, slot) * B;
#pragma unroll
        for (int b = 0; b < B; ++b) v[b] = (double)src[b];
    }
    // Segmented inclusive scan inside runs of equal column.
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const int32_t ocol = __shfl_up(col, d, G);
        double ov[B];
#pragma unroll
        for (int b = 0; b < B; ++b) ov[b] = __shfl_up(v[b], d, G);
        if (lane >= d && ocol == col) {
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] += ov[b];
        }
    }
    const int32_t ncol = __shfl_down(col, 1, G);
    const bool tail = valid && (lane == n - 1 || ncol != col);
    if (tail) {
        const uint64_t le = wl == 63 ? ~0ull : ((1ull << (wl + 1)) - 1);
        const int32_t rank = __popcll(gheads & le) - 1;
        for (int dr = 0; dr < dpn; ++dr) {
            const int64_t rs = (int64_t)row_ptr[a * dpn + dr] + (int64_t)rank * dpn;
            for (int dc = 0; dc < dpn; ++dc) {
                col_ind[rs + dc] = col * dpn + dc;
                val[rs + dc] = v[dr * dpn + dc];
            }
        }
    }
}

// ------------------------------------------- wave-per-row (LDS) rows ---
__device__ __forceinline__ int32_t wave_excl_scan(int32_t x, int32_t* total) {
    const int lane = threadIdx.x & 63;
    int32_t s = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t o = __shfl_up(s, d, 64);
        if (lane >= d) s += o;
    }
    *total = __shfl(s, 63, 64);
    return s - x;
}

// One 64-thread block (one wave) per row at a time; `cap` keys of LDS.
template <int B, typename VT, bool WRITE, typename RP>
__global__ __launch_bounds__(64) void k_rows_big(ListPool p, const int32_t* list, int32_t* u,
                                                 const RP* row_ptr, int32_t* col_ind, double* val,
                                                 int dpn, int cap) {
    extern __shared__ uint64_t smem_big[];
    uint64_t* sk = smem_big;
    uint16_t* ss = reinterpret_cast<uint16_t*>(sk + cap);
    const int lane = threadIdx.x;
    const int64_t count = list ? (int64_t)p.misc[MISC_BIG_COUNT] : p.n_nodes;
    for (int64_t r = blockIdx.x; r < count; r += gridDim.x) {
        const int64_t a = list ? (int64_t)list[r] : r;
        const int32_t* h = p.hdr + a * ALLS_DIR;
        const int32_t n = min(h[0], ALLS_DIR * p.C);
        if (n > cap) {
            if (lane == 0) atomicMax(&p.misc[MISC_ERR + ERR_ROW_TOO_BIG], n);
            continue;
        }
        int P = 1;
        while (P < n) P <<= 1;
        for (int i = lane; i < P; i += 64) {
            sk[i] = i < n ? pack_key(p.keys[entry_index(p, a, h, i)]) : KEY_INF;
            ss[i] = (uint16_t)i;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const uint64_t ka = sk[i], kb = sk[l];
                        const bool up = (i & k) == 0;
                        if (up ? (ka > kb) : (ka < kb)) {
                            sk[i] = kb;
                            sk[l] = ka;
                            const uint16_t t = ss[i];
                            ss[i] = ss[l];
                            ss[l] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // Each lane owns a contiguous segment of sorted positions.
        const int seg = (P + 63) / 64;
        const int lo = lane * seg;
        const int hi = min(lo + seg, n);
        int32_t nh = 0;
        for (int i = lo; i < hi; ++i)
            nh += (i == 0 || key_col(sk[i]) != key_col(sk[i - 1])) ? 1 : 0;
        int32_t total;
        const int32_t base = wave_excl_scan(nh, &total);
        if (!WRITE) {
            if (lane == 0) u[a] = total;
        } else {
            int32_t rank = base;
            for (int i = lo; i < hi; ++i) {
                const int32_t col = key_col(sk[i]);
                if (!(i == 0 || col != key_col(sk[i - 1]))) continue;
                double v[B];
#pragma unroll
                for (int b = 0; b < B; ++b) v[b] = 0.0;
                for (int j = i; j < n && key_col(sk[j]) == col; ++j) {
                    const VT* src = static_cast<const VT*>(p.vals) + entry_index(p, a, h, ss[j]) * B;
#pragma unroll
                    for (int b = 0; b < B; ++b) v[b] += (double)src[b];
                }
                for (int dr = 0; dr < dpn; ++dr) {
                    const int64_t rs = (int64_t)row_ptr[a * dpn + dr] + (int64_t)rank * dpn;
                    for (int dc = 0; dc < dpn; ++dc) {
                        col_ind[rs + dc] = col * dpn + dc;
                        val[rs + dc] = v[dr * dpn + dc];
                    }
                }
                ++rank;
            }
        }
        __syncthreads();
    }
}

// Unique-column count of long rows with an LDS hash set (one 64-thread block
// per row at a time, `ts` table slots, ts >= 2 * ALLS_DIR * C).  Counting
// needs no order, so no sort: each entry inserts its column with one LDS
// compare-and-swap (linear probing).
__global__ __launch_bounds__(64) void k_rows_hash_count(ListPool p, const int32_t* list, int32_t* u, int ts) {
    extern __shared__ int32_t table[];
    const int lane = threadIdx.x;
    const int64_t count = list ? (int64_t)p.misc[MISC_BIG_COUNT] : p.n_nodes;
    const int shift = 32 - (31 - __clz(ts));
    for (int64_t r = blockIdx.x; r < count; r += gridDim.x) {
        const int64_t a = list ? (int64_t)list[r] : r;
        const int32_t* h = p.hdr + a * ALLS_DIR;
        const int32_t n = min(h[0], ALLS_DIR * p.C);
        for (int i = lane; i < ts; i += 64) table[i] = -1;
        __syncthreads();
        int32_t mine = 0;
        for (int i = lane; i < n; i += 64) {
            const int32_t col = p.keys[entry_index(p, a, h, i)].x;
            uint32_t slot = ((uint32_t)col * 0x9E3779B1u) >> shift;
            for (int probe = 0; probe < ts; ++probe) {
                const int32_t old = atomicCAS(&table[slot], -1, col);
                if (old == -1) {
                    ++mine;
                    break;
                }
                if (old == col) break;
                slot = (slot + 1) & (uint32_t)(ts - 1);
            }
        }
        int32_t total;
        wave_excl_scan(mine, &total);
        if (lane == 0) u[a] = total;
        __syncthreads();
    }
}

// ------------------------------------------------------------- scan ---
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER;

template <typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* total) {
    __shared__ T wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    T s = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = __shfl_up(s, d, 64);
        if (lane >= d) s += o;
    }
    if (lane == 63) wsum[w] = s;
    __syncthreads();
    T off = 0, tot = 0;
    for (int i = 0; i < nw; ++i) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + s - x;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const int32_t* u, int64_t n, int64_t* tiles) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS)
        if (base + i < n) s += u[base + i];
    int64_t tot;
    block_excl_scan<int64_t>(s, &tot);
    if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// Exclusive scan of the tile sums in place; tiles[ntiles] = grand total.
__global__ __launch_bounds__(1024) void k_scan_top(int64_t* tiles, int64_t ntiles) {
    int64_t carry = 0;
    for (int64_t s0 = 0; s0 < ntiles; s0 += 1024) {
        const int64_t i = s0 + threadIdx.x;
        const int64_t v = i < ntiles ? tiles[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan<int64_t>(v, &tot);
        if (i < ntiles) tiles[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) tiles[ntiles] = carry;
}

// row_ptr[a*dpn+dr] = dpn*(off_a*dpn + dr*u_a), off_a = sum of u over nodes < a.
template <typename RP>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int32_t* u, int64_t n, const int64_t* tiles,
                                                             int dpn, RP* row_ptr) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int32_t loc[SCAN_PER];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        loc[i] = base + i < n ? u[base + i] : 0;
        s += loc[i];
    }
    int64_t tot;
    int64_t off = tiles[blockIdx.x] + block_excl_scan<int64_t>(s, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        const int64_t a = base + i;
        if (a < n) {
            for (int dr = 0; dr < dpn; ++dr)
                row_ptr[a * dpn + dr] = (RP)(dpn * (off * dpn + (int64_t)dr * loc[i]));
            off += loc[i];
            if (a == n - 1) row_ptr[n * dpn] = (RP)(off * dpn * dpn);
        }
    }
}

template <int B, typename VT, bool WRITE, typename RP>
void rows_dispatch(const ListPool& p, int G, const int32_t* big_list, int32_t* u, int32_t* bl,
                   const RP* rp, int32_t* ci, double* v, int dpn, hipStream_t s) {
    if (G > 0) {
        const int64_t threads = p.n_nodes * G;
        const unsigned blocks = (unsigned)((threads + 255) / 256);
        if (G == 16)
            hipLaunchKernelGGL((k_rows_small<16, B, VT, WRITE, RP>), dim3(blocks), dim3(256), 0, s, p, u, bl, rp, ci, v, dpn);
        else if (G == 32)
            hipLaunchKernelGGL((k_rows_small<32, B, VT, WRITE, RP>), dim3(blocks), dim3(256), 0, s, p, u, bl, rp, ci, v, dpn);
        else
            hipLaunchKernelGGL((k_rows_small<64, B, VT, WRITE, RP>), dim3(blocks), dim3(256), 0, s, p, u, bl, rp, ci, v, dpn);
    }
    // Long rows: over the big-row list (G > 0) or over every row (G == 0).
    int cap = 1;
    while (cap < ALLS_DIR * p.C) cap <<= 1;
    int64_t grid = G > 0 ? 2048 : p.n_nodes;
    if (grid > 262144) grid = 262144;
    if (grid < 1) grid = 1;
    if (!WRITE) {
        const int ts = 2 * cap;
        hipLaunchKernelGGL(k_rows_hash_count, dim3((unsigned)grid), dim3(64), (size_t)ts * sizeof(int32_t), s, p,
                           G > 0 ? big_list : nullptr, u, ts);
        return;
    }
    const size_t lds = (size_t)cap * (sizeof(uint64_t) + sizeof(uint16_t));
    hipLaunchKernelGGL((k_rows_big<B, VT, WRITE, RP>), dim3((unsigned)grid), dim3(64), lds, s, p,
                       G > 0 ? big_list : nullptr, u, rp, ci, v, dpn, cap);
}

}  // namespace

namespace alls {

int64_t scan_tile_count(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_count(const ListPool& p, int G, int f32vals, int32_t* u, int32_t* big_list, hipStream_t s) {
    if (p.B == 4)
        rows_dispatch<4, double, false, int32_t>(p, G, big_list, u, big_list, nullptr, nullptr, nullptr, 2, s);
    else if (f32vals)
        rows_dispatch<9, float, false, int32_t>(p, G, big_list, u, big_list, nullptr, nullptr, nullptr, 3, s);
    else
        rows_dispatch<9, double, false, int32_t>(p, G, big_list, u, big_list, nullptr, nullptr, nullptr, 3, s);
}

void launch_scan_total(const int32_t* u, int64_t n, int64_t* tiles, int64_t ntiles, hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tiles, ntiles);
}

void launch_scan_apply(const int32_t* u, int64_t n, const int64_t* tiles, int dpn, void* row_ptr, int i64,
                       hipStream_t s) {
    const int64_t ntiles = scan_tile_count(n);
    if (ntiles == 0) {
        // No nodes: row_ptr = {0}.
        if (i64) hipMemsetAsync(row_ptr, 0, sizeof(int64_t), s);
        else hipMemsetAsync(row_ptr, 0, sizeof(int32_t), s);
        return;
    }
    if (i64)
        hipLaunchKernelGGL(k_scan_apply<int64_t>, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles, dpn,
                           static_cast<int64_t*>(row_ptr));
    else
        hipLaunchKernelGGL(k_scan_apply<int32_t>, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles, dpn,
                           static_cast<int32_t*>(row_ptr));
}

void launch_write(const ListPool& p, int G, int f32vals, int dpn, const int32_t* big_list, const void* row_ptr,
                  int i64, int32_t* col_ind, double* val, hipStream_t s) {
    (void)dpn;
    if (i64) {
        const int64_t* rp = static_cast<const int64_t*>(row_ptr);
        if (p.B == 4) rows_dispatch<4, double, true, int64_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 2, s);
        else if (f32vals) rows_dispatch<9, float, true, int64_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 3, s);
        else rows_dispatch<9, double, true, int64_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 3, s);
    } else {
        const int32_t* rp = static_cast<const int32_t*>(row_ptr);
        if (p.B == 4) rows_dispatch<4, double, true, int32_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 2, s);
        else if (f32vals) rows_dispatch<9, float, true, int32_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 3, s);
        else rows_dispatch<9, double, true, int32_t>(p, G, big_list, nullptr, nullptr, rp, col_ind, val, 3, s);
    }
}

}  // namespace alls
