// consolidate.hip -- A-LLS lists -> CSR on the device (gfx950).
//
// The reference builds row_ptr/col_ind on the host with std::set
// (Acc:605-645) and its convertALLStoCSRKernel (Acc:471-501) walks each row's
// list with one thread, binary-searching every node's column in the host
// pattern.  Here the pattern itself is produced on the device:
//   count : per node row, count its unique columns u[a];
//   scan  : u -> row_ptr (tile sums, one-block scan of the tile sums, apply);
//   write : merge equal columns (sum in element-id order) and write
//           col_ind / val of the dpn DOF rows of the node.
// Rows of <= G entries use G lanes and a register bitonic sort through
// cross-lane shuffles (32-bit columns to count; (column, element) keys to
// write, after which the head lane of each run adds the run in order).
// Longer rows (Hex20, or unstructured outliers) use a wave per row and LDS: a
// hash set of the columns for the count; for the write, the same hash set,
// column ranks by counting, a bucket per column and an element-id sort inside
// each bucket (a handful of entries), instead of a full sort of the row; rows
// beyond 1024 entries use an LDS bitonic sort (k_rows_big).  Either way every
// column's contributions are summed one by one in ascending element id from
// 0, so the values do not depend on the order in which atomics appended the
// entries.
#include "alls_internal.h"

namespace {

constexpr uint64_t KEY_INF = ~0ull;

__device__ __forceinline__ uint64_t pack_key(int2 kv) {
    return ((uint64_t)(uint32_t)kv.x << 32) | (uint32_t)kv.y;
}
__device__ __forceinline__ int32_t key_col(uint64_t k) { return (int32_t)(k >> 32); }

// Cross-lane moves inside 16-lane rows by DPP (a VALU operand modifier, no
// LDS round trip), with ds_bpermute (__shfl_*) beyond them.  The sorts and
// run sums below call these with constant distances once unrolled.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
// Value of lane (lane ^ j) of the same G-lane group.
__device__ __forceinline__ uint32_t xor_u32(uint32_t v, int j, int G) {
    switch (j) {
        case 1: return dpp_u32<0xB1>(v);                   // quad_perm [1,0,3,2]
        case 2: return dpp_u32<0x4E>(v);                   // quad_perm [2,3,0,1]
        case 4: return dpp_u32<0x1B>(dpp_u32<0x141>(v));   // row_half_mirror, then quad_perm [3,2,1,0]
        case 8: return dpp_u32<0x128>(v);                  // row_ror:8
        default: return (uint32_t)__shfl_xor((int)v, j, G);
    }
}
__device__ __forceinline__ uint64_t xor_u64(uint64_t v, int j, int G) {
    return ((uint64_t)xor_u32((uint32_t)(v >> 32), j, G) << 32) | xor_u32((uint32_t)v, j, G);
}
// Value of lane (lane + d) of the same group, d >= 1; garbage where lane + d
// leaves the group (callers mask those lanes).  DPP row_shl for G == 16.
template <int G>
__device__ __forceinline__ uint32_t down_u32(uint32_t v, int d) {
    if constexpr (G == 16) {
        switch (d) {
            case 1: return dpp_u32<0x101>(v);
            case 2: return dpp_u32<0x102>(v);
            case 3: return dpp_u32<0x103>(v);
            case 4: return dpp_u32<0x104>(v);
            case 5: return dpp_u32<0x105>(v);
            case 6: return dpp_u32<0x106>(v);
            case 7: return dpp_u32<0x107>(v);
            case 8: return dpp_u32<0x108>(v);
            case 9: return dpp_u32<0x109>(v);
            case 10: return dpp_u32<0x10A>(v);
            case 11: return dpp_u32<0x10B>(v);
            case 12: return dpp_u32<0x10C>(v);
            case 13: return dpp_u32<0x10D>(v);
            case 14: return dpp_u32<0x10E>(v);
            default: return dpp_u32<0x10F>(v);
        }
    }
    const int lane = (int)(threadIdx.x & (G - 1));
    return (uint32_t)__shfl((int)v, lane + d, G);
}
// Value of lane (lane - 1) of the same group (lane 0: garbage).
template <int G>
__device__ __forceinline__ uint32_t up1_u32(uint32_t v) {
    if constexpr (G == 16) return dpp_u32<0x111>(v);   // row_shr:1
    return (uint32_t)__shfl_up((int)v, 1, G);
}
template <int G>
__device__ __forceinline__ double down_f64(double x, int d) {
    const uint64_t b = __double_as_longlong(x);
    const uint64_t r = ((uint64_t)down_u32<G>((uint32_t)(b >> 32), d) << 32) | down_u32<G>((uint32_t)b, d);
    return __longlong_as_double(r);
}

__device__ __forceinline__ int64_t entry_index(const ListPool& p, int64_t a, const int32_t* h, int32_t s) {
    // Callers clamp s below ALLS_DIR * C.  A chunk that was never published
    // (only after a directory overflow, which the host then retries) maps to
    // the home chunk so that every access stays in bounds.
    const int32_t k = chunk_of(p, s);
    const int32_t d = k == 0 ? 0 : h[k];
    const int64_t ch = d > 0 ? p.n_nodes + (int64_t)(d - 1) : a;
    return ch * p.C + (s - k * p.C);
}

// List layout as a compile-time NG: 4 = Q4 groups (element id per group of 4
// slots, in geid), 3 = T3 line groups (key block inside each group's 128-byte
// line), 0 = per-entry keys.  gkey (Q4 groups): the key of slot r given its
// group's element e: the column node is conn[e][r % 4].  An element id out of range (a group left unwritten after
// a failed chunk wait, which the host then retries) reads as column a with
// *pad set (its value counts as zero), so no access leaves the arrays.
template <int NG>
__device__ __forceinline__ int32_t grp(int32_t r) {
    return NG == 4 ? (r >> 2) : (int32_t)(__umulhi((uint32_t)r, 0xAAAAAAABu) >> 1);
}
template <int NG>
__device__ __forceinline__ int2 gkey(const ListPool& p, int64_t a, int32_t r, int32_t e, bool* pad) {
    const int32_t q = r - grp<NG>(r) * NG;
    const bool bad = (uint64_t)(uint32_t)e >= (uint64_t)p.n_elems;
    const int32_t b = bad ? -1 : p.conn[(int64_t)e * p.npe + q];
    *pad = b < 0;
    return make_int2(b < 0 ? (int32_t)a : b, bad ? 0 : e);
}

// Chunk and in-chunk slot of slot s of row a (as entry_index).
__device__ __forceinline__ void entry_loc(const ListPool& p, int64_t a, const int32_t* h, int32_t s, int64_t* ch,
                                          int32_t* r) {
    const int32_t k = chunk_of(p, s);
    const int32_t d = k == 0 ? 0 : h[k];
    *ch = d > 0 ? p.n_nodes + (int64_t)(d - 1) : a;
    *r = s - k * p.C;
}

// 2D lists: writes the (column, element) keys of a row that the wave-per-row
// kernels will take (more than G entries), from the group element ids, and
// zeroes the value of each unused triangle slot (read as column a).  Lanes
// 0..G-1 of the row's group take slots lane, lane+G, ...
template <int NG, int B, typename VT>
__device__ __forceinline__ void materialize_keys(const ListPool& p, int64_t a, int lane, int G, int32_t n) {
    const int32_t* h = p.hdr + a * ALLS_DIR;
    for (int32_t sl = lane; sl < n; sl += G) {
        int64_t ch;
        int32_t r;
        entry_loc(p, a, h, sl, &ch, &r);
        bool pad;
        const int2 key = gkey<NG>(p, a, r, p.geid[ch * p.gpc + grp<NG>(r)], &pad);
        const int64_t idx = ch * p.C + r;
        p.keys[idx] = key;
        if (pad) {
            VT* v = static_cast<VT*>(p.vals) + idx * B;
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] = VT(0);
        }
    }
}

// entry_index with the row's directory already in registers (wave-uniform,
// loaded once per row): no memory access per entry.
struct RowDir {
    int32_t d[ALLS_DIR];
    __device__ __forceinline__ void load(const int32_t* h) {
        const int4 x = reinterpret_cast<const int4*>(h)[0], y = reinterpret_cast<const int4*>(h)[1];
        d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w;
        d[4] = y.x; d[5] = y.y; d[6] = y.z; d[7] = y.w;
    }
    __device__ __forceinline__ int64_t entry(const ListPool& p, int64_t a, int32_t s) const {
        const int32_t k = chunk_of(p, s);
        int32_t dk = 0;
#pragma unroll
        for (int j = 1; j < ALLS_DIR; ++j) dk = k == j ? d[j] : dk;
        const int64_t ch = dk > 0 ? p.n_nodes + (int64_t)(dk - 1) : a;
        return ch * p.C + (s - k * p.C);
    }
};

// ------------------------------------------------ register (small) rows ---
// One node row of at most G entries, merged in registers by its G lanes:
// the keys and values are loaded (home chunk together with the header),
// sorted by (column, element) with a bitonic network of shuffles, and the
// head lane of each run of equal column adds the run one by one in element
// order (0 + v1 + v2 + ..., the serial order, so this path gives the same
// bits as every other path of the library).  Returns false for a longer row.
template <int G, int B, typename VT>
struct SmallRow {
    int32_t col, rank, u;
    bool head;
    double v[B];
};

// What a lane loads of a row before anything depends on it: its entry of
// the home chunk (key and, for the write pass, value block) and the row
// header's count.  Rows are loaded a few at a time before any is merged, so
// that more loads are in flight per wave.
template <int B, typename VT>
struct SmallLoad {
    int2 kh;
    int32_t n;
    VT vh[B];
};

// T3 line groups: key and value-block index of in-chunk slot r of chunk ch.
// The column and the element are two loads from computed addresses (no
// select on loaded data), so that callers can issue the loads of several
// rows before waiting for any.
__device__ __forceinline__ int2 line3_key(const ListPool& p, int64_t ch, int32_t r, int64_t* vslot) {
    const int32_t g = (int32_t)(__umulhi((uint32_t)r, 0xAAAAAAABu) >> 1), q = r - 3 * g;
    const int64_t line = ch * p.gpc + g;
    const int32_t* kb = reinterpret_cast<const int32_t*>(static_cast<const char*>(p.vals) + line * 128 + 96);
    *vslot = line * LINE3_SLOTS + q;
    return make_int2(kb[1 + q], kb[0]);
}
__device__ __forceinline__ int32_t line3_col(const ListPool& p, int64_t ch, int32_t r) {
    const int32_t g = (int32_t)(__umulhi((uint32_t)r, 0xAAAAAAABu) >> 1), q = r - 3 * g;
    const int64_t line = ch * p.gpc + g;
    return reinterpret_cast<const int32_t*>(static_cast<const char*>(p.vals) + line * 128 + 96)[1 + q];
}

template <int NG, int B, typename VT, bool VALS>
__device__ __forceinline__ void small_row_load(const ListPool& p, int64_t a, int lane, SmallLoad<B, VT>& L) {
    L.kh = make_int2(0, 0);
    if (lane < p.C) {
        int64_t vs = a * p.C + lane;
        if constexpr (NG == 4) L.kh.y = p.geid[a * p.gpc + grp<NG>(lane)];   // group layout: the group's element
        else if constexpr (NG == 3) L.kh = line3_key(p, a, lane, &vs);
        else L.kh = p.keys[vs];
        if constexpr (VALS) {
            const VT* src = static_cast<const VT*>(p.vals) + vs * B;
#pragma unroll
            for (int b = 0; b < B; ++b) L.vh[b] = src[b];
        }
    }
    L.n = p.hdr[a * ALLS_DIR];
}

// Second step after the loads: the row's entry count (clamped to the
// directory), the entries of overflow chunks, and for 2D lists each slot's
// column from its group's element.  Rows are resolved a few at a time before
// any is merged, so that these dependent loads overlap too.  A row longer
// than G is left alone (the wave-per-row kernels take it).
template <int G, int NG, int B, typename VT>
__device__ __forceinline__ void small_row_keys(const ListPool& p, int64_t a, int lane, SmallLoad<B, VT>& L) {
    const int32_t n = min(L.n, ALLS_DIR * p.C);
    L.n = n;
    if (n > G) return;
    if (lane >= p.C && lane < n) {   // overflow chunk (rare for rows this short)
        int64_t ch;
        int32_t r;
        entry_loc(p, a, p.hdr + a * ALLS_DIR, lane, &ch, &r);
        int64_t ei = ch * p.C + r;
        if constexpr (NG == 4) L.kh.y = p.geid[ch * p.gpc + grp<NG>(r)];
        else if constexpr (NG == 3) L.kh = line3_key(p, ch, r, &ei);
        else L.kh = p.keys[ei];
        const VT* src = static_cast<const VT*>(p.vals) + ei * B;
#pragma unroll
        for (int b = 0; b < B; ++b) L.vh[b] = src[b];
    }
    if (NG == 4 && lane < n) {   // group layout: column from the group's element
        bool pad;
        L.kh = gkey<NG>(p, a, lane, L.kh.y, &pad);
        if (pad) {
#pragma unroll
            for (int b = 0; b < B; ++b) L.vh[b] = VT(0);
        }
    }
}

// After small_row_keys.  Returns false for a row longer than G.  `stage`
// holds B values per thread of the block (the caller's LDS): each lane parks
// its entry's values there, the keys are sorted with their original slots,
// and the head lane of each run of equal columns reads the run's values by
// slot and adds them one by one, 0 + v1 + v2 + ... (the serial order).
template <int G, int B, typename VT>
__device__ __forceinline__ bool small_row_merge(const ListPool& p, int64_t a, int lane, SmallLoad<B, VT>& L,
                                                SmallRow<G, B, VT>& r, VT* stage) {
    const int32_t n = L.n;
    if (n > G) return false;
    const int2 kh = L.kh;
    const int wl = threadIdx.x & 63;
    const int gbase = wl - lane;
#pragma unroll
    for (int b = 0; b < B; ++b) stage[threadIdx.x * B + b] = L.vh[b];
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << (G & 63)) - 1) << gbase);
    uint64_t key = KEY_INF;
    int32_t slot = lane;
    if (lane < n) key = pack_key(kh);
    // Sorted by (column, element, slot): slots are distinct, so the order is
    // total and the same on every run.  Equal (column, element) keys come
    // from an element that lists a node twice; their slots follow the
    // element's (row node, column node) order, the serial order.
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t ok = xor_u64(key, j, G);
            const int32_t os = (int32_t)xor_u32((uint32_t)slot, j, G);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            const bool less = ok < key || (ok == key && os < slot);
            const bool take = (lower == up) ? less : !less;
            if (take) {
                key = ok;
                slot = os;
            }
        }
    }
    r.col = key_col(key);
    const bool valid = lane < n;
    const int32_t pcol = (int32_t)up1_u32<G>((uint32_t)(key >> 32));
    r.head = valid && (lane == 0 || pcol != r.col);
    const uint64_t gheads = __ballot(r.head) & gmask;
    r.u = __popcll(gheads);
    const uint64_t le = wl == 63 ? ~0ull : ((1ull << (wl + 1)) - 1);
    r.rank = __popcll(gheads & le) - 1;
    const uint64_t above = wl == 63 ? 0ull : (gheads & (~0ull << (wl + 1)));
    const int32_t run = r.head ? (above ? (__ffsll((unsigned long long)above) - 1 - gbase) : n) - lane : 0;
    const VT* gstage = stage + (threadIdx.x - lane) * B;   // the group's values, by original slot
#pragma unroll
    for (int b = 0; b < B; ++b) r.v[b] = 0.0;
    auto step = [&](int32_t sl) {
#pragma unroll
        for (int b = 0; b < B; ++b) r.v[b] += (double)gstage[sl * B + b];
    };
    if (run > 0) step(slot);
    // Slots of the run's other members: sorted positions lane+1, lane+2, ...
    // Unrolled for 16-lane rows (DPP needs constant distances).
    if constexpr (G == 16) {
#pragma unroll
        for (int d = 1; d < G; ++d) {
            if (!__any(d < run)) break;
            const int32_t sd = (int32_t)down_u32<G>((uint32_t)slot, d);
            if (d < run) step(sd);
        }
    } else {
        for (int d = 1; d < G; ++d) {
            if (!__any(d < run)) break;
            const int32_t sd = (int32_t)down_u32<G>((uint32_t)slot, d);
            if (d < run) step(sd);
        }
    }
    return true;
}

// small_row_merge for T3 line groups in 32-lane rows: the same sums with a
// 32-bit sort key.  Every group of 3 slots is one element row (slots 3g ..
// 3g+2, columns node 0, 1, 2 of the element), so (column, element, slot)
// order is (column, rank of the group by (element, group), slot in group):
// key = column << 5 | (group rank * 3 + slot in group), for columns below
// 2^27 (the host keys every entry otherwise).  The values are parked in LDS
// at that element-order slot, which the low 5 bits of a sorted key name.
// The group ranks take one pass over the row's (at most 10) groups; the
// 64-bit compare-exchange of (key, slot) pairs it replaces was most of the
// pass's VALU work.
template <int G, int B, typename VT>
__device__ __forceinline__ bool small_row_merge_line3(int lane, SmallLoad<B, VT>& L, SmallRow<G, B, VT>& r,
                                                      VT* stage) {
    static_assert(G == 32, "line groups take 32-lane rows");
    const int32_t n = L.n;
    if (n > G) return false;
    const int wl = threadIdx.x & 63;
    const int gbase = wl - lane;
    const bool valid = lane < n;
    const int32_t g = (lane * 11) >> 5;   // lane / 3 for lanes 0..31
    const int32_t q = lane - 3 * g;
    const uint32_t e = (uint32_t)L.kh.y;
    const int32_t ngr = (n + 2) / 3;
    int32_t grank = 0;
#pragma unroll
    for (int gg = 0; gg < (G + 2) / 3; ++gg) {
        if (!__any(gg < ngr)) break;
        const uint32_t eo = (uint32_t)__shfl((int)e, gg * 3, G);
        grank += (gg < ngr && (eo < e || (eo == e && gg < g))) ? 1 : 0;
    }
    const int32_t es = min(grank * 3 + q, G - 1);   // < n for every valid lane of a whole-group row
    uint32_t key = valid ? (((uint32_t)L.kh.x << 5) | (uint32_t)es) : 0xffffffffu;
    VT* gstage = stage + (threadIdx.x - lane) * B;   // the group's values, by element-order slot
    if (valid) {
#pragma unroll
        for (int b = 0; b < B; ++b) gstage[es * B + b] = L.vh[b];
    }
    // Distinct keys (valid lanes) sort first; the unused lanes' all-ones
    // keys stay behind them.  Compare-exchange as min / max and a select
    // (a conditional on the lane's direction compiles to divergent branches).
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t ok = xor_u32(key, j, G);
            const bool keep_low = ((lane & k) == 0) == ((lane & j) == 0);
            key = keep_low ? min(key, ok) : max(key, ok);
        }
    }
    const uint64_t gmask = ((1ull << G) - 1) << gbase;
    r.col = (int32_t)(key >> 5);
    const uint32_t pk = up1_u32<G>(key);
    r.head = valid && (lane == 0 || (pk >> 5) != (key >> 5));
    const uint64_t gheads = __ballot(r.head) & gmask;
    r.u = __popcll(gheads);
    const uint64_t le = wl == 63 ? ~0ull : ((1ull << (wl + 1)) - 1);
    r.rank = __popcll(gheads & le) - 1;
    const uint64_t above = wl == 63 ? 0ull : (gheads & (~0ull << (wl + 1)));
    const int32_t run = r.head ? (above ? (__ffsll((unsigned long long)above) - 1 - gbase) : n) - lane : 0;
#pragma unroll
    for (int b = 0; b < B; ++b) r.v[b] = 0.0;
    auto step = [&](int32_t sl) {
#pragma unroll
        for (int b = 0; b < B; ++b) r.v[b] += (double)gstage[sl * B + b];
    };
    if (run > 0) step((int32_t)(key & 31));
    for (int d = 1; d < G; ++d) {
        if (!__any(d < run)) break;
        const int32_t sd = (int32_t)(down_u32<G>(key, d) & 31);
        if (d < run) step(sd);
    }
    return true;
}

// Head lanes write their dpn x dpn block at rank `rank` of the node's rows,
// whose first entries are rs[dr].
template <int B, typename OT>
__device__ __forceinline__ void small_row_store(int32_t col, int32_t rank, const double* v, const int64_t* rs,
                                                int32_t* col_ind, OT* val, int dpn) {
    if constexpr (B == 4) {
        // 2 DOF per node: row starts are even, so each block row is one
        // aligned int2 + double2 store.
#pragma unroll
        for (int dr = 0; dr < 2; ++dr) {
            const int64_t q = rs[dr] + (int64_t)rank * 2;
            *reinterpret_cast<int2*>(col_ind + q) = make_int2(col * 2, col * 2 + 1);
            *reinterpret_cast<double2*>(val + q) = make_double2(v[dr * 2], v[dr * 2 + 1]);
        }
    } else {
        for (int dr = 0; dr < dpn; ++dr) {
            const int64_t q = rs[dr] + (int64_t)rank * dpn;
            for (int dc = 0; dc < dpn; ++dc) {
                col_ind[q + dc] = col * dpn + dc;
                val[q + dc] = (OT)v[dr * dpn + dc];
            }
        }
    }
}

// G lanes per row; a wave takes R rounds of 64/G consecutive rows, and the
// loads of all R rounds are issued before the first row is merged.  Row of
// round r: wave base + r * 64/G + group, so that each round reads and writes
// one contiguous stretch as a single round would.
template <int G, int R, int NG, int B, typename VT, bool WRITE, typename RP, typename OT>
__global__ __launch_bounds__(256) void k_rows_small(ListPool p, int32_t* u, int32_t* big_list,
                                                    const RP* row_ptr, int32_t* col_ind, OT* val,
                                                    int dpn) {
    constexpr int GPW = 64 / G;   // row groups per wave
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(gid % G);
    const int64_t wbase = gid / 64 * (GPW * R) + (gid % 64) / G;
    if (wbase >= p.n_nodes) return;
    if constexpr (!WRITE) {
        // Counting needs only the columns: a bitonic sort of 32-bit keys.
        // kc: the entry's column (keyed lists) or its group's element (2D).
        int32_t kc[R], hn[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t a = wbase + r * GPW;
            kc[r] = 0;
            hn[r] = 0;
            if (a < p.n_nodes) {
                if (lane < p.C) {   // with the header, not after it
                    if constexpr (NG == 4) kc[r] = p.geid[a * p.gpc + grp<NG>(lane)];
                    else if constexpr (NG == 3) kc[r] = line3_col(p, a, lane);
                    else kc[r] = p.keys[a * p.C + lane].x;
                }
                hn[r] = min(p.hdr[a * ALLS_DIR], ALLS_DIR * p.C);
            }
        }
        // Overflow entries and 2D columns of every round, before any sort.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t a = wbase + r * GPW;
            const int32_t n = hn[r];
            if (a >= p.n_nodes || n > G) continue;
            if (lane >= p.C && lane < n) {
                int64_t ch;
                int32_t rr;
                entry_loc(p, a, p.hdr + a * ALLS_DIR, lane, &ch, &rr);
                if constexpr (NG == 4) kc[r] = p.geid[ch * p.gpc + grp<NG>(rr)];
                else if constexpr (NG == 3) kc[r] = line3_col(p, ch, rr);
                else kc[r] = p.keys[ch * p.C + rr].x;
            }
            if (NG == 4 && lane < n) {
                bool pad;
                kc[r] = gkey<NG>(p, a, lane, kc[r], &pad).x;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t a = wbase + r * GPW;
            if (a >= p.n_nodes) break;
            const int32_t n = hn[r];
            if (n > G) {
                if (lane == 0) big_list[atomicAdd(&p.misc[MISC_BIG_COUNT], 1)] = (int32_t)a;
                if (NG == 4) materialize_keys<NG, B, VT>(p, a, lane, G, n);
                continue;
            }
            uint32_t c = lane < n ? (uint32_t)kc[r] : 0xffffffffu;
#pragma unroll
            for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const uint32_t oc = xor_u32(c, j, G);
                    const bool keep_low = ((lane & k) == 0) == ((lane & j) == 0);
                    c = keep_low ? min(c, oc) : max(c, oc);
                }
            }
            const uint32_t pc = up1_u32<G>(c);
            const int wl = threadIdx.x & 63;
            const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << (wl - lane));
            const uint64_t heads = __ballot(lane < n && (lane == 0 || pc != c));
            if (lane == 0) u[a] = __popcll(heads & gmask);
        }
    } else {
        __shared__ VT vstage[256 * B];
        SmallLoad<B, VT> L[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (wbase + r * GPW < p.n_nodes) small_row_load<NG, B, VT, true>(p, wbase + r * GPW, lane, L[r]);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (wbase + r * GPW < p.n_nodes) small_row_keys<G, NG, B, VT>(p, wbase + r * GPW, lane, L[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t a = wbase + r * GPW;
            if (a >= p.n_nodes) break;
            SmallRow<G, B, VT> row;
            bool fits;
            if constexpr (NG == 3 && G == 32) fits = small_row_merge_line3<G, B, VT>(lane, L[r], row, vstage);
            else fits = small_row_merge<G, B, VT>(p, a, lane, L[r], row, vstage);
            if (!fits) continue;   // long row: other kernel
            if (row.head) {
                int64_t rs[3];
                for (int dr = 0; dr < dpn; ++dr) rs[dr] = (int64_t)row_ptr[a * dpn + dr];
                small_row_store<B, OT>(row.col, row.rank, row.v, rs, col_ind, val, dpn);
            }
        }
    }
}

// Count pass for Q4 lists (groups of 4 slots): a lane per group, 4 lanes
// (a DPP quad) per row, so a wave counts 16 rows.  Each lane loads its
// group's element id and the element's 4 nodes (one 16-byte load); a column
// counts if no lower lane of the quad holds it and no earlier column of its
// own lane equals it, and the quad adds its counts.  Rows of more than 16
// entries go to the wave-per-row kernels as in k_rows_small.
template <int B, typename VT>
__global__ __launch_bounds__(256) void k_count_quad(ListPool p, int32_t* u, int32_t* big_list) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a = gid >> 2;
    const int g = (int)(gid & 3);
    if (a >= p.n_nodes) return;   // whole quads
    const int32_t* h = p.hdr + a * ALLS_DIR;
    int32_t e = 4 * g < p.C ? p.geid[a * p.gpc + g] : -1;   // with the header, not after it
    const int32_t n = min(h[0], ALLS_DIR * p.C);
    if (n > 16) {
        if (g == 0) big_list[atomicAdd(&p.misc[MISC_BIG_COUNT], 1)] = (int32_t)a;
        materialize_keys<4, B, VT>(p, a, g, 4, n);
        return;
    }
    const bool valid = 4 * g < n;
    if (valid && 4 * g >= p.C) {   // overflow chunk (chunks of fewer than 16 slots)
        int64_t ch;
        int32_t r;
        entry_loc(p, a, h, 4 * g, &ch, &r);
        e = p.geid[ch * p.gpc + grp<4>(r)];
    }
    int32_t v[4] = {-1, -1, -1, -1};
    if (valid) {
        if ((uint64_t)(uint32_t)e < (uint64_t)p.n_elems) {
            const int4 c = *reinterpret_cast<const int4*>(p.conn + (int64_t)e * 4);
            v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w;
        } else {   // a group left unwritten (the host retries): column a, as gkey reads it
            v[0] = v[1] = v[2] = v[3] = (int32_t)a;
        }
    }
    // Columns of the other lanes of the quad (DPP quad broadcasts).
    int32_t w[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        w[0][j] = (int32_t)dpp_u32<0x00>((uint32_t)v[j]);   // quad_perm [0,0,0,0]
        w[1][j] = (int32_t)dpp_u32<0x55>((uint32_t)v[j]);   // [1,1,1,1]
        w[2][j] = (int32_t)dpp_u32<0xAA>((uint32_t)v[j]);   // [2,2,2,2]
        w[3][j] = (int32_t)dpp_u32<0xFF>((uint32_t)v[j]);   // [3,3,3,3]
    }
    int32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        bool first = v[i] >= 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) first &= !(k < g && w[k][j] == v[i]);
#pragma unroll
        for (int j = 0; j < i; ++j) first &= v[j] != v[i];
        cnt += first ? 1 : 0;
    }
    cnt += (int32_t)dpp_u32<0xB1>((uint32_t)cnt);   // quad_perm [1,0,3,2]
    cnt += (int32_t)dpp_u32<0x4E>((uint32_t)cnt);   // quad_perm [2,3,0,1]
    if (g == 0) u[a] = cnt;
}

// Count, scan and write of small rows in ONE pass (2D meshes whose rows all
// fit G lanes).  Blocks take tiles of fused_iter(G) * 256/G rows in the order
// they start (ticket from an atomic counter), merge their rows as the write
// pass does (results kept in registers), and find the rows' CSR offsets with
// a decoupled look-back over the tiles started before them: each tile
// publishes its aggregate, then its inclusive prefix once known; one wave
// reads 64 predecessor states at a time, adds the aggregates back to the
// nearest published prefix, and moves the window back by 64 while there is
// none.  Every tile it waits on has started, so the waits end; they are
// bounded anyway (spin limit -> error flag; the host then runs the separate
// passes).  Saves the count pass: the keys are read and sorted once.
// The tile states are relaxed atomics: each state word carries its own
// payload, so no ordering of other memory is needed, and acquire/release at
// device scope would write back or invalidate the XCD's caches per access.
constexpr int64_t TILE_AGG = int64_t(1) << 62, TILE_PREFIX = int64_t(2) << 62;
constexpr int64_t TILE_VALUE = TILE_AGG - 1;
// Rows per lane group in a tile: 6 for 32-lane rows (48 rows per tile;
// measured against 2, 4 and 8: T3 1.38 -> 1.35 ms, mixed 1.11 -> 1.04 ms at
// 1500x1500), 4 for 16-lane rows (a tile's rows must fit one wave's scan).
__host__ __device__ constexpr int fused_iter(int G) { return G == 32 ? 6 : 4; }

__device__ __forceinline__ int64_t wave_sum64(int64_t x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

template <int G, int NG, typename RP>
__global__ __launch_bounds__(256) void k_rows_fused(ListPool p, int64_t* tstat, int64_t n_tiles, RP* row_ptr,
                                                    int32_t* col_ind, double* val) {
    constexpr int RPI = 256 / G;                 // rows per iteration
    constexpr int FUSED_ITER = fused_iter(G);
    constexpr int RPT = RPI * FUSED_ITER;        // rows per tile (<= 64)
    constexpr int B = 4, D = 2;
    static_assert(RPT <= 64, "one wave scans a tile's rows");
    __shared__ int32_t s_tile;
    __shared__ int64_t s_u[RPT];
    __shared__ int64_t s_base;
    if (threadIdx.x == 0) s_tile = atomicAdd(&p.misc[MISC_TICKET], 1);
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= n_tiles) return;   // never: the grid has exactly n_tiles blocks
    const int rib = threadIdx.x / G, lane = threadIdx.x % G;
    SmallRow<G, B, double> r[FUSED_ITER];
    SmallLoad<B, double> L[FUSED_ITER];
    __shared__ double vstage[256 * B];
#pragma unroll
    for (int it = 0; it < FUSED_ITER; ++it) {
        const int64_t a = tile * RPT + it * RPI + rib;
        if (a < p.n_nodes) small_row_load<NG, B, double, true>(p, a, lane, L[it]);
    }
#pragma unroll
    for (int it = 0; it < FUSED_ITER; ++it) {
        const int64_t a = tile * RPT + it * RPI + rib;
        if (a < p.n_nodes) small_row_keys<G, NG, B, double>(p, a, lane, L[it]);
    }
#pragma unroll
    for (int it = 0; it < FUSED_ITER; ++it) {
        const int64_t a = tile * RPT + it * RPI + rib;
        r[it].u = 0;
        r[it].head = false;
        bool fits = true;
        if (a < p.n_nodes) {
            if constexpr (NG == 3 && G == 32) fits = small_row_merge_line3<G, B, double>(lane, L[it], r[it], vstage);
            else fits = small_row_merge<G, B, double>(p, a, lane, L[it], r[it], vstage);
        }
        if (!fits) {
            // the host checked that no row is longer than G; never taken
            if (lane == 0) atomicMax(&p.misc[MISC_ERR + ERR_ROW_TOO_BIG], G + 1);
            r[it].u = 0;
            r[it].head = false;
        }
        if (lane == 0) s_u[it * RPI + rib] = a < p.n_nodes ? r[it].u : 0;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int wl = threadIdx.x;
        // Exclusive scan of the tile's row counts.
        const int64_t x = wl < RPT ? s_u[wl] : 0;
        int64_t incl = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t o = __shfl_up(incl, d, 64);
            if (wl >= d) incl += o;
        }
        const int64_t agg = __shfl(incl, 63, 64);
        if (wl < RPT) s_u[wl] = incl - x;
        if (wl == 0)
            __hip_atomic_store(&tstat[tile], (tile == 0 ? TILE_PREFIX : TILE_AGG) | agg, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        int64_t excl = 0;
        for (int64_t j = tile - 1; j >= 0; j -= 64) {
            const int64_t idx = j - wl;
            int64_t st = idx >= 0 ? __hip_atomic_load(&tstat[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : TILE_PREFIX;
            for (int32_t it = 0;; ++it) {
                const bool pending = (st & ~TILE_VALUE) == 0;
                if (!__any(pending)) break;
                if (it >= ALLS_SPIN_MAX) {
                    if (wl == 0) st_agent(&p.misc[MISC_ERR + ERR_LOOKBACK], 1);
                    st = TILE_PREFIX;   // give up: the host discards this result
                    continue;
                }
                __builtin_amdgcn_s_sleep(1);
                if (pending) st = __hip_atomic_load(&tstat[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const uint64_t pm = __ballot((st & TILE_PREFIX) != 0);
            if (pm) {   // nearest predecessor with a prefix: stop there
                const int first = __ffsll((unsigned long long)pm) - 1;
                excl += wave_sum64(wl <= first ? (st & TILE_VALUE) : 0);
                break;
            }
            excl += wave_sum64(st & TILE_VALUE);
        }
        if (wl == 0) {
            if (tile > 0)
                __hip_atomic_store(&tstat[tile], TILE_PREFIX | (excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            s_base = excl;
            if (tile == n_tiles - 1) {   // the last tile knows the total
                const int64_t total = excl + agg;
                row_ptr[p.n_nodes * D] = (RP)(total * D * D);
                st_agent(&p.misc[MISC_TOTAL_LO], (int32_t)(total & 0x7fffffff));
                st_agent(&p.misc[MISC_TOTAL_HI], (int32_t)(total >> 31));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < FUSED_ITER; ++it) {
        const int64_t a = tile * RPT + it * RPI + rib;
        if (a >= p.n_nodes) continue;
        const int64_t off = s_base + s_u[it * RPI + rib];
        int64_t rs[D];
#pragma unroll
        for (int dr = 0; dr < D; ++dr) rs[dr] = D * (off * D + (int64_t)dr * r[it].u);
        if (lane == 0) {
#pragma unroll
            for (int dr = 0; dr < D; ++dr) row_ptr[a * D + dr] = (RP)rs[dr];
        }
        if (r[it].head) small_row_store<B, double>(r[it].col, r[it].rank, r[it].v, rs, col_ind, val, D);
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

// Rows beyond k_rows_merge's LDS (more than MERGE_CAP_MAX entries): one
// 64-thread block (one wave) per row at a time, a bitonic sort of `cap` keys
// staged in LDS.
template <int B, typename VT, bool WRITE, typename RP, typename OT>
__global__ __launch_bounds__(64) void k_rows_big(ListPool p, const int32_t* list, int32_t* u,
                                                 const RP* row_ptr, int32_t* col_ind, OT* val,
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
            sk[i] = i < n ? pack_key(list_key(p, entry_index(p, a, h, i))) : KEY_INF;
            ss[i] = (uint16_t)i;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const uint64_t ka = sk[i], kb = sk[l];
                        const uint16_t sa = ss[i], sb = ss[l];
                        const bool up = (i & k) == 0;
                        // (key, slot) order: total, slots are distinct.
                        const bool a_gt = ka > kb || (ka == kb && sa > sb);
                        if (up ? a_gt : !a_gt) {
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
                    const VT* src = static_cast<const VT*>(p.vals) + list_val_slot(p, entry_index(p, a, h, ss[j])) * B;
#pragma unroll
                    for (int b = 0; b < B; ++b) v[b] += (double)src[b];
                }
                for (int dr = 0; dr < dpn; ++dr) {
                    const int64_t rs = (int64_t)row_ptr[a * dpn + dr] + (int64_t)rank * dpn;
                    for (int dc = 0; dc < dpn; ++dc) {
                        col_ind[rs + dc] = col * dpn + dc;
                        val[rs + dc] = (OT)v[dr * dpn + dc];
                    }
                }
                ++rank;
            }
        }
        __syncthreads();
    }
}

// Hash slot of column `col` in a table of 2^(32-shift) slots.
__device__ __forceinline__ uint32_t col_hash(int32_t col, int shift) {
    return ((uint32_t)col * 0x9E3779B1u) >> shift;
}

// Write pass of long rows without sorting them.  One 64-thread block (one
// wave) per row at a time, `cap` (power of two, >= the longest row) sizes
// the LDS arrays:
//   1. load the row's keys, insert each column into a hash set (2*cap slots);
//   2. number the occupied slots (unique index), which also gives u;
//   3. count entries per unique column;
//   4. rank of each column = number of smaller columns (counting over u);
//   5. bucket offsets by rank (scan), entries dropped into their bucket;
//   6. lane per column: insertion-sort its bucket by element id, sum the
//      values in that order and write the dpn x dpn block; lanes take
//      consecutive columns, so the stores are contiguous.
template <int B, typename VT, typename RP, typename OT>
__global__ __launch_bounds__(64) void k_rows_merge(ListPool p, const int32_t* list, const RP* row_ptr,
                                                   int32_t* col_ind, OT* val, int cap) {
    constexpr int D = B == 4 ? 2 : 3;
    extern __shared__ int32_t smem_merge[];
    const int ts = 2 * cap;
    int32_t* table = smem_merge;       // [ts]   column, then unique index of the slot
    int32_t* eeid = table + ts;        // [cap]  element id of entry i
    int32_t* eu = eeid + cap;          // [cap]  table slot, then unique index of entry i
    int32_t* ucol = eu + cap;          // [cap]  column of unique index q (padded to 4)
    int32_t* ucnt = ucol + cap;        // [cap]  entries of unique index q
    int32_t* urank = ucnt + cap;       // [cap]  rank of unique index q
    int32_t* rcol = urank + cap;       // [cap]  column of rank r
    int32_t* goff = rcol + cap;        // [cap+1] bucket start of rank r
    int32_t* gfill = goff + cap + 1;   // [cap]  bucket fill of rank r
    int32_t* memb = gfill + cap;       // [cap]  entries grouped by rank
    const int lane = threadIdx.x;
    const int shift = 32 - (31 - __clz(ts));
    const int64_t count = list ? (int64_t)p.misc[MISC_BIG_COUNT] : p.n_nodes;
    for (int64_t row = xcd_block(blockIdx.x, gridDim.x); row < count; row += gridDim.x) {
        const int64_t a = list ? (int64_t)list[row] : row;
        const int32_t* h = p.hdr + a * ALLS_DIR;
        const int32_t n = min(h[0], ALLS_DIR * p.C);
        if (n > cap) {   // the host sizes cap from the count pass; never taken
            if (lane == 0) atomicMax(&p.misc[MISC_ERR + ERR_ROW_TOO_BIG], n);
            continue;
        }
        for (int i = lane; i < ts; i += 64) table[i] = -1;
        for (int i = lane; i < cap; i += 64) {
            ucnt[i] = 0;
            gfill[i] = 0;
        }
        __syncthreads();
        // 1. keys -> hash set; eu[i] = slot of the entry's column.
        for (int i = lane; i < n; i += 64) {
            const int2 k = list_key(p, entry_index(p, a, h, i));
            eeid[i] = k.y;
            uint32_t slot = col_hash(k.x, shift);
            for (int probe = 0; probe < ts; ++probe) {
                const int32_t old = atomicCAS(&table[slot], -1, k.x);
                if (old == -1 || old == k.x) break;
                slot = (slot + 1) & (uint32_t)(ts - 1);
            }
            eu[i] = (int32_t)slot;
        }
        __syncthreads();
        // 2. unique index = order of the occupied slot; the slot keeps it.
        const int tseg = ts / 64;
        int32_t occ = 0;
        for (int j = 0; j < tseg; ++j) occ += table[lane * tseg + j] >= 0 ? 1 : 0;
        int32_t u;
        int32_t q = wave_excl_scan(occ, &u);
        for (int j = 0; j < tseg; ++j) {
            const int32_t c = table[lane * tseg + j];
            if (c >= 0) {
                ucol[q] = c;
                table[lane * tseg + j] = q++;
            }
        }
        const int u4 = (u + 3) & ~3;
        if (lane < u4 - u) ucol[u + lane] = 0x7fffffff;
        __syncthreads();
        // 3. entries per unique column.
        for (int i = lane; i < n; i += 64) {
            const int32_t qi = table[eu[i]];
            eu[i] = qi;
            atomicAdd(&ucnt[qi], 1);
        }
        __syncthreads();
        // 4. ranks.
        for (int qi = lane; qi < u; qi += 64) {
            const int32_t c = ucol[qi];
            int32_t r = 0;
            for (int k = 0; k < u4; k += 4) {
                const int4 w = *reinterpret_cast<const int4*>(ucol + k);
                r += (w.x < c) + (w.y < c) + (w.z < c) + (w.w < c);
            }
            urank[qi] = r;
            rcol[r] = c;
            goff[r] = ucnt[qi];
        }
        __syncthreads();
        // 5. bucket offsets (exclusive scan over ranks), then the buckets.
        {
            const int rseg = (u + 63) / 64;
            const int lo = min(lane * rseg, u), hi = min(lo + rseg, u);
            int32_t sum = 0;
            for (int r = lo; r < hi; ++r) sum += goff[r];
            int32_t tot;
            int32_t off = wave_excl_scan(sum, &tot);
            for (int r = lo; r < hi; ++r) {
                const int32_t c = goff[r];
                goff[r] = off;
                off += c;
            }
            if (lane == 0) goff[u] = n;
        }
        __syncthreads();
        for (int i = lane; i < n; i += 64) {
            const int32_t r = urank[eu[i]];
            memb[goff[r] + atomicAdd(&gfill[r], 1)] = i;
        }
        __syncthreads();
        // 6. one lane per column.
        for (int r = lane; r < u; r += 64) {
            const int32_t g0 = goff[r], g1 = goff[r + 1];
            for (int m = g0 + 1; m < g1; ++m) {
                const int32_t x = memb[m];
                const int32_t ex = eeid[x];
                int k = m - 1;
                while (k >= g0 && (eeid[memb[k]] > ex || (eeid[memb[k]] == ex && memb[k] > x))) {
                    memb[k + 1] = memb[k];
                    --k;
                }
                memb[k + 1] = x;
            }
            double v[B];
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] = 0.0;
            for (int m = g0; m < g1; ++m) {
                const VT* src = static_cast<const VT*>(p.vals) + list_val_slot(p, entry_index(p, a, h, memb[m])) * B;
#pragma unroll
                for (int b = 0; b < B; ++b) v[b] += (double)src[b];
            }
            const int32_t col = rcol[r];
#pragma unroll
            for (int dr = 0; dr < D; ++dr) {
                const int64_t rs = (int64_t)row_ptr[a * D + dr] + (int64_t)r * D;
#pragma unroll
                for (int dc = 0; dc < D; ++dc) {
                    col_ind[rs + dc] = col * D + dc;
                    val[rs + dc] = (OT)v[dr * D + dc];
                }
            }
        }
        __syncthreads();
    }
}

// k_rows_merge for rows of at most 64*EPL entries (EPL = entries per lane),
// with the memory latency taken off the critical path: every lane issues
// the loads of its entries' keys AND values at the top of the row, keeps
// the values in registers while the LDS phases (hash set, ranks, buckets,
// element-id sort) run, and then hands them over through LDS, one DOF row
// (dpn values per entry) at a time, in bucket order.  One global-memory
// round trip per row instead of one per bucket member.
template <int B, typename VT, typename RP, int EPL, typename OT>
__global__ __launch_bounds__(64) void k_rows_merge_reg(ListPool p, const int32_t* list, const RP* row_ptr,
                                                       int32_t* col_ind, OT* val, int count_slot) {
    constexpr int D = B == 4 ? 2 : 3;
    constexpr int CAP = 64 * EPL;
    constexpr int TS = 2 * CAP <= 128 ? 128 : 2 * CAP <= 256 ? 256 : 2 * CAP <= 512 ? 512 : 1024;   // power of two
    constexpr int PH = TS + 6 * CAP;   // phase arrays
    __shared__ __align__(16) int32_t sm[PH + 3 * CAP + 1];
    // Phase arrays; the first PH ints are dead once the buckets are sorted
    // and then hold the value stage (CAP * D doubles).
    int32_t* table = sm;               // [TS]
    int32_t* eeid = table + TS;        // [CAP]
    int32_t* memb = eeid + CAP;        // [CAP]
    int32_t* ucol = memb + CAP;        // [CAP]
    int32_t* ucnt = ucol + CAP;        // [CAP]
    int32_t* urank = ucnt + CAP;       // [CAP]
    int32_t* gfill = urank + CAP;      // [CAP]
    double* stage = reinterpret_cast<double*>(sm);
    int32_t* eu = sm + PH;             // [CAP]  slot, unique index, then bucket position of entry i
    int32_t* goff = eu + CAP;          // [CAP+1]
    int32_t* rcol = goff + CAP + 1;    // [CAP]
    static_assert(CAP * D * 2 <= PH, "value stage must fit the dead phase arrays");
    const int lane = threadIdx.x;
    constexpr int shift = 32 - (31 - __builtin_clz(TS));
    const int64_t count = list ? (int64_t)p.misc[count_slot] : p.n_nodes;
    for (int64_t row = xcd_block(blockIdx.x, gridDim.x); row < count; row += gridDim.x) {
        const int64_t a = list ? (int64_t)list[row] : row;
        // Home-chunk slots (< C) are loaded before the header arrives: they
        // do not depend on it.  Slots past the count are loaded but unused.
        int2 kk[EPL];
        VT vv[EPL][B];
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int i = lane + 64 * k;
            kk[k] = make_int2(0, 0);
#pragma unroll
            for (int b = 0; b < B; ++b) vv[k][b] = VT(0);
            if (i < p.C) {
                const int64_t ei = a * p.C + i;
                kk[k] = list_key(p, ei);
                const VT* src = static_cast<const VT*>(p.vals) + list_val_slot(p, ei) * B;
#pragma unroll
                for (int b = 0; b < B; ++b) vv[k][b] = src[b];
            }
        }
        RowDir dir;
        dir.load(p.hdr + a * ALLS_DIR);
        const int32_t n = min(dir.d[0], ALLS_DIR * p.C);
        if (n > CAP) {   // the host picks EPL from the count pass; never taken
            if (lane == 0) atomicMax(&p.misc[MISC_ERR + ERR_ROW_TOO_BIG], n);
            continue;
        }
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int i = lane + 64 * k;
            if (i >= p.C && i < n) {   // overflow chunks
                const int64_t ei = dir.entry(p, a, i);
                kk[k] = list_key(p, ei);
                const VT* src = static_cast<const VT*>(p.vals) + list_val_slot(p, ei) * B;
#pragma unroll
                for (int b = 0; b < B; ++b) vv[k][b] = src[b];
            }
        }
        for (int i = lane; i < TS; i += 64) table[i] = -1;
        for (int i = lane; i < CAP; i += 64) {
            ucnt[i] = 0;
            gfill[i] = 0;
        }
        __syncthreads();
        // Hash set of the columns; eu[i] = slot of the entry's column.
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                eeid[i] = kk[k].y;
                uint32_t slot = col_hash(kk[k].x, shift);
                for (int probe = 0; probe < TS; ++probe) {
                    const int32_t old = atomicCAS(&table[slot], -1, kk[k].x);
                    if (old == -1 || old == kk[k].x) break;
                    slot = (slot + 1) & (uint32_t)(TS - 1);
                }
                eu[i] = (int32_t)slot;
            }
        }
        __syncthreads();
        // Unique index = order of the occupied slot.
        constexpr int TSEG = TS / 64;
        int32_t occ = 0;
#pragma unroll
        for (int j = 0; j < TSEG; ++j) occ += table[lane * TSEG + j] >= 0 ? 1 : 0;
        int32_t u;
        int32_t q = wave_excl_scan(occ, &u);
#pragma unroll
        for (int j = 0; j < TSEG; ++j) {
            const int32_t c = table[lane * TSEG + j];
            if (c >= 0) {
                ucol[q] = c;
                table[lane * TSEG + j] = q++;
            }
        }
        const int u4 = (u + 3) & ~3;
        if (lane < u4 - u) ucol[u + lane] = 0x7fffffff;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                const int32_t qi = table[eu[i]];
                eu[i] = qi;
                atomicAdd(&ucnt[qi], 1);
            }
        }
        __syncthreads();
        // Ranks by counting smaller columns.
        for (int qi = lane; qi < u; qi += 64) {
            const int32_t c = ucol[qi];
            int32_t r = 0;
            for (int k = 0; k < u4; k += 4) {
                const int4 w = *reinterpret_cast<const int4*>(ucol + k);
                r += (w.x < c) + (w.y < c) + (w.z < c) + (w.w < c);
            }
            urank[qi] = r;
            rcol[r] = c;
            goff[r] = ucnt[qi];
        }
        __syncthreads();
        {
            const int rseg = (u + 63) / 64;
            const int lo = min(lane * rseg, u), hi = min(lo + rseg, u);
            int32_t sum = 0;
            for (int r = lo; r < hi; ++r) sum += goff[r];
            int32_t tot;
            int32_t off = wave_excl_scan(sum, &tot);
            for (int r = lo; r < hi; ++r) {
                const int32_t c = goff[r];
                goff[r] = off;
                off += c;
            }
            if (lane == 0) goff[u] = n;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                const int32_t r = urank[eu[i]];
                memb[goff[r] + atomicAdd(&gfill[r], 1)] = i;
            }
        }
        __syncthreads();
        // Element-id order inside each bucket; eu[i] = position of entry i.
        for (int r = lane; r < u; r += 64) {
            const int32_t g0 = goff[r], g1 = goff[r + 1];
            for (int m = g0 + 1; m < g1; ++m) {
                const int32_t x = memb[m];
                const int32_t ex = eeid[x];
                int k = m - 1;
                while (k >= g0 && (eeid[memb[k]] > ex || (eeid[memb[k]] == ex && memb[k] > x))) {
                    memb[k + 1] = memb[k];
                    --k;
                }
                memb[k + 1] = x;
            }
            for (int m = g0; m < g1; ++m) eu[memb[m]] = m;
        }
        __syncthreads();
        // One DOF row at a time: stage the entries' dpn values at their
        // bucket positions, then a lane per column sums its bucket in order.
#pragma unroll
        for (int dr = 0; dr < D; ++dr) {
#pragma unroll
            for (int k = 0; k < EPL; ++k) {
                const int i = lane + 64 * k;
                if (i < n) {
                    const int32_t pos = eu[i];
#pragma unroll
                    for (int dc = 0; dc < D; ++dc) stage[pos * D + dc] = (double)vv[k][dr * D + dc];
                }
            }
            __syncthreads();
            const int64_t rbase = (int64_t)row_ptr[a * D + dr];
            for (int r = lane; r < u; r += 64) {
                double acc[D];
#pragma unroll
                for (int dc = 0; dc < D; ++dc) acc[dc] = 0.0;
                for (int m = goff[r]; m < goff[r + 1]; ++m) {
#pragma unroll
                    for (int dc = 0; dc < D; ++dc) acc[dc] += stage[m * D + dc];
                }
                const int64_t rs = rbase + (int64_t)r * D;
                const int32_t col = rcol[r];
#pragma unroll
                for (int dc = 0; dc < D; ++dc) {
                    col_ind[rs + dc] = col * D + dc;
                    val[rs + dc] = (OT)acc[dc];
                }
            }
            __syncthreads();
        }
    }
}

// Reads lane `l`'s copy of a 64-bit value (l uniform).
__device__ __forceinline__ int64_t readlane64(int64_t x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Rows of 3D lists whose elements reserve SEG slots at a time (Hex20: one
// reservation of 20 per element and row), so entry i lies in segment i / SEG
// and a segment is one element's row strip with SEG distinct columns.  One
// wave per row; entries sit SPS = 64 / SEG whole segments to a register slot
// (entry i = 60 k + lane for SEG = 20, lanes 60..63 idle), K slots, so a
// segment lies in one slot k = s / SPS.  Then:
//   1. LDS hash set of the columns -> unique index q (ballot order);
//   2. segments ranked by element id;
//   3. FP64 accumulators acc[q][9] in LDS, zeroed, then the segments added
//      one at a time in element-id order (one register slot, one lane group
//      per step): inside a step the columns are distinct, so each
//      accumulator sees 0 + v1 + v2 + ... in element order, the serial
//      order of every other path (same bits);
//   4. column ranks by counting (16 columns per LDS round trip);
//   5. a lane per column writes its 3 x 3 block into the 3 DOF rows.
// Loads: a row's keys, then its values, then the next row's header, all
// issued together and unconditionally (lanes past the row read its slot 0),
// with the header and the row starts spread over lanes and read back with
// readlane where used.  A load under a branch, or a uniform load (moved to a
// scalar register at once), would make the compiler wait for every load
// issued before it.  Segment lists (SG) have no keys: each lane's strip
// element id is loaded with the row's header (it does not depend on it) and
// its column is read from the connectivity, so the columns are as many
// round trips away as keys are.
// Rows outside that shape (a length that is not a multiple of SEG or above
// the slots, a segment with mixed element ids or a repeated column, two
// segments of one element, more than UMAX columns) are listed in `defer`
// (count misc[MISC_DEFER]) for k_rows_merge_reg.
constexpr int SEG_UMAX = 96;
template <typename VT, typename RP, int K, int SEG, typename OT, bool SG>
__global__ __launch_bounds__(64, 4) void k_rows_merge_seg(ListPool p, const RP* row_ptr, int32_t* col_ind,
                                                          OT* val, int32_t* defer) {
    constexpr int B = 9, D = 3;
    constexpr int SPS = 64 / SEG;        // segments per register slot
    constexpr int LPS = SPS * SEG;       // lanes per slot
    constexpr int CAP = LPS * K;
    constexpr int NS = SPS * K;
    constexpr int TS = 256;              // hash slots (a row of more than UMAX columns is deferred)
    constexpr int UMAX = SEG_UMAX;
    static_assert(NS <= 32 && UMAX % 16 == 0 && UMAX <= 128, "segment masks, rank padding, two columns per lane");
    static_assert(TS * sizeof(int32_t) <= UMAX * B * sizeof(double), "the hash table lives in the accumulators");
    // The accumulators hold the hash table until every entry has its q.
    __shared__ __align__(16) double acc[UMAX * B];
    __shared__ __align__(16) int32_t ucol[UMAX];
    __shared__ int32_t rq[UMAX];
    __shared__ uint32_t smask[UMAX];
    __shared__ int32_t seid[NS];
    int32_t* table = reinterpret_cast<int32_t*>(acc);
    const int lane = threadIdx.x;
    const int sl = lane / SEG;           // segment of this lane inside a slot (SPS for idle lanes)
    constexpr int shift = 32 - (31 - __builtin_clz(TS));
    int64_t a = xcd_block(blockIdx.x, gridDim.x);
    // Header of the next row: lane l holds word l % ALLS_DIR.
    int32_t hv = p.hdr[min(a, p.n_nodes - 1) * ALLS_DIR + (lane & (ALLS_DIR - 1))];
    for (; a < p.n_nodes; a += gridDim.x) {
        // SG: the strip element id of each register slot's entry in the
        // home chunk (strip clamped into it; entries past it are read by the
        // overflow path), issued before the header is waited for.
        int32_t gw[K];
        if constexpr (SG) {
#pragma unroll
            for (int k = 0; k < K; ++k) gw[k] = p.geid[a * p.gpc + min(lane < LPS ? SPS * k + sl : 0, p.gpc - 1)];
        }
        const int32_t hd = hv;   // this row's header
        const int32_t n = min(__builtin_amdgcn_readlane(hd, 0), ALLS_DIR * p.C);
        const int64_t an = a + gridDim.x;
        const int lim = min(n, p.C);
        int2 kk[K];
        VT vv[K][B];
        if constexpr (!SG) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = LPS * k + lane;
                kk[k] = p.keys[a * p.C + (lane < LPS && i < lim ? i : 0)];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = LPS * k + lane;
            const VT* src = static_cast<const VT*>(p.vals) + (a * p.C + (lane < LPS && i < lim ? i : 0)) * B;
#pragma unroll
            for (int b = 0; b < B; ++b) vv[k][b] = src[b];
        }
        if constexpr (SG) {
            // After the values: these wait for the strip ids, the values
            // only for the header.
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = LPS * k + lane;
                const int32_t e = seg_elem(p, gw[k]);
                kk[k] = make_int2(p.conn[(int64_t)e * p.npe + (lane < LPS && i < lim ? lane - SEG * sl : 0)], e);
            }
        }
        // Starts of the 3 DOF rows (lane dr holds row dr's).
        const int64_t rv = (int64_t)row_ptr[a * D + min(lane, D - 1)];
        hv = p.hdr[min(an, p.n_nodes - 1) * ALLS_DIR + (lane & (ALLS_DIR - 1))];
        if (n > p.C && n <= CAP) {   // overflow chunks (rows longer than the chunk)
            RowDir dir;
#pragma unroll
            for (int j = 0; j < ALLS_DIR; ++j) dir.d[j] = __builtin_amdgcn_readlane(hd, j);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = LPS * k + lane;
                if (lane < LPS && i >= p.C && i < n) {
                    const int64_t ei = dir.entry(p, a, i);
                    kk[k] = SG ? list_key(p, ei) : p.keys[ei];
                    const VT* src = static_cast<const VT*>(p.vals) + ei * B;
#pragma unroll
                    for (int b = 0; b < B; ++b) vv[k][b] = src[b];
                }
            }
            // Settle this rare path's loads here: otherwise the compiler's
            // wait for the keys, merged over both paths, would cover the
            // common path's value loads too.
            __builtin_amdgcn_s_waitcnt(0);
        }
        const int ns = n <= CAP ? n / SEG : 0;   // segments (rows over CAP: none, deferred)
        bool ok = ns * SEG == n;
        reinterpret_cast<int4*>(table)[lane] = make_int4(-1, -1, -1, -1);   // TS = 256
        for (int i = lane; i < UMAX; i += 64) smask[i] = 0u;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = LPS * k + lane;
            if (lane < LPS && i < n && lane % SEG == 0) seid[SPS * k + sl] = kk[k].y;
        }
        __syncthreads();
        // 1. Hash set of the columns.  The K slots probe together: one LDS
        // round trip per probe step for all of them.
        uint32_t slot[K];
        bool pend[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = LPS * k + lane;
            pend[k] = ok && lane < LPS && i < n;
            slot[k] = pend[k] ? col_hash(kk[k].x, shift) : 0u;
        }
        for (int probe = 0; probe < TS; ++probe) {
            int32_t old[K];
#pragma unroll
            for (int k = 0; k < K; ++k) old[k] = pend[k] ? atomicCAS(&table[slot[k]], -1, kk[k].x) : -1;
            bool more = false;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (pend[k] && old[k] != -1 && old[k] != kk[k].x) slot[k] = (slot[k] + 1) & (uint32_t)(TS - 1);
                else pend[k] = false;
                more |= pend[k];
            }
            if (__ballot(more) == 0) break;
        }
        __syncthreads();
        // Unique index: occupied slots numbered in (j, lane) order.
        int32_t tv[TS / 64];
#pragma unroll
        for (int j = 0; j < TS / 64; ++j) tv[j] = table[j * 64 + lane];
        int32_t u = 0;
#pragma unroll
        for (int j = 0; j < TS / 64; ++j) {
            const int32_t c = tv[j];
            const uint64_t m = __ballot(c >= 0);
            const int32_t q = u + (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (c >= 0) {
                if (q < UMAX) ucol[q] = c;
                table[j * 64 + lane] = q;
            }
            u += __popcll(m);
        }
        ok = ok && u <= UMAX;
        __syncthreads();
        // Each entry's q; element ids constant inside a segment; no column
        // twice in one segment.
        int32_t eq[K];
        bool bad = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = LPS * k + lane;
            eq[k] = 0;
            if (ok && lane < LPS && i < n) {
                const int32_t q = table[slot[k]];
                eq[k] = q;
                const int sg = SPS * k + sl;
                bad |= seid[sg] != kk[k].y;
                if (q < UMAX) bad |= (atomicOr(&smask[q], 1u << sg) >> sg) & 1u;
            }
        }
        // 2. Segment ranks by element id (distinct ids required); lane r of
        // `order` holds the segment of rank r.
        const int32_t my_eid = lane < ns ? seid[lane] : 0;
        int32_t rk = 0, same = 0;
        for (int sg = 0; sg < ns; ++sg) {
            const int32_t e = __builtin_amdgcn_readlane(my_eid, sg);
            rk += e < my_eid ? 1 : 0;
            same += e == my_eid ? 1 : 0;
        }
        bad |= lane < ns && same != 1;
        ok = ok && __ballot(bad) == 0;
        if (!ok) {
            if (lane == 0) defer[atomicAdd(&p.misc[MISC_DEFER], 1)] = (int32_t)a;
            __syncthreads();
            continue;
        }
        if (lane < ns) smask[rk] = (uint32_t)lane;   // segment masks are done with
        const int u16 = (u + 15) & ~15;
        if (lane < u16 - u) ucol[u + lane] = 0x7fffffff;
        __syncthreads();   // table reads done: the accumulators take its place
        for (int x = lane; 2 * x < u * B; x += 64) reinterpret_cast<double2*>(acc)[x] = make_double2(0.0, 0.0);
        const int32_t order = lane < ns ? (int32_t)smask[lane] : 0;
        __syncthreads();
        // 3. Segments in element-id order: one slot and lane group per step.
        for (int sr = 0; sr < ns; ++sr) {
            const int sg = __builtin_amdgcn_readlane(order, sr);
            const int k = sg / SPS;
            if (sl == sg - SPS * k) {
#pragma unroll
                for (int kk2 = 0; kk2 < K; ++kk2) {
                    if (kk2 == k) {   // uniform
                        double* d = acc + eq[kk2] * B;
#pragma unroll
                        for (int b = 0; b < B; ++b) d[b] += (double)vv[kk2][b];
                    }
                }
            }
        }
        // 4. Column ranks: lanes take q = lane, and q = lane + 64 when the
        // row has more than 64 columns (uniform branch: most rows have fewer).
        {
            const int32_t c0 = lane < u ? ucol[lane] : 0x7fffffff;
            int32_t r0 = 0;
            if (u <= 64) {
                for (int k = 0; k < u16; k += 16) {
                    int4 w[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) w[t] = *reinterpret_cast<const int4*>(ucol + k + 4 * t);
#pragma unroll
                    for (int t = 0; t < 4; ++t) r0 += (w[t].x < c0) + (w[t].y < c0) + (w[t].z < c0) + (w[t].w < c0);
                }
            } else {
                const int32_t c1 = lane + 64 < u ? ucol[lane + 64] : 0x7fffffff;
                int32_t r1 = 0;
                for (int k = 0; k < u16; k += 16) {
                    int4 w[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) w[t] = *reinterpret_cast<const int4*>(ucol + k + 4 * t);
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        r0 += (w[t].x < c0) + (w[t].y < c0) + (w[t].z < c0) + (w[t].w < c0);
                        r1 += (w[t].x < c1) + (w[t].y < c1) + (w[t].z < c1) + (w[t].w < c1);
                    }
                }
                if (lane + 64 < u) rq[r1] = lane + 64;
            }
            if (lane < u) rq[r0] = lane;
        }
        __syncthreads();
        // 5. A lane per column: its 3 x 3 block into the 3 DOF rows.
        int64_t rbase[D];
#pragma unroll
        for (int dr = 0; dr < D; ++dr) rbase[dr] = readlane64(rv, dr);
        for (int r = lane; r < u; r += 64) {
            const int32_t q = rq[r];
            const int32_t col = ucol[q];
            double w[B];
#pragma unroll
            for (int b = 0; b < B; ++b) w[b] = acc[q * B + b];
#pragma unroll
            for (int dr = 0; dr < D; ++dr) {
                const int64_t rs = rbase[dr] + (int64_t)r * D;
#pragma unroll
                for (int dc = 0; dc < D; ++dc) {
                    col_ind[rs + dc] = col * D + dc;
                    val[rs + dc] = (OT)w[dr * D + dc];
                }
            }
        }
        __syncthreads();
    }
}

// LDS bytes of k_rows_merge for `cap`.
constexpr size_t merge_lds(int cap) { return (size_t)(2 * cap + 9 * cap + 1) * sizeof(int32_t); }
constexpr int MERGE_CAP_MAX = 1024;   // 45 KB of LDS
constexpr int LONG_CAP_MAX = 4096;    // k_rows_big: 40 KB of LDS

// Unique-column count of long rows with an LDS hash set (one 64-thread block
// per row at a time, `ts` table slots, ts >= 2 * ALLS_DIR * C).  Counting
// needs no order, so no sort: each entry inserts its column with one LDS
// compare-and-swap (linear probing).
// With `defer` set, rows that need more than `ts` slots are appended to it
// (count in MISC_BIG_COUNT) for a second launch with a larger table, so the
// common short rows run with a small LDS allocation and more waves per CU.
__global__ __launch_bounds__(64) void k_rows_hash_count(ListPool p, const int32_t* list, int32_t* u, int ts,
                                                        int32_t* defer) {
    extern __shared__ int32_t table[];
    const int lane = threadIdx.x;
    const int64_t count = list ? (int64_t)p.misc[MISC_BIG_COUNT] : p.n_nodes;
    int32_t longest = 0;   // this block's longest row; one atomic per block
    for (int64_t r = blockIdx.x; r < count; r += gridDim.x) {
        const int64_t a = list ? (int64_t)list[r] : r;
        // The first two keys of this lane in the home chunk are loaded
        // together with the row header (they do not depend on it).
        int32_t hk0 = 0, hk1 = 0;
        if (lane < p.C) hk0 = list_key(p, a * p.C + lane).x;
        if (lane + 64 < p.C) hk1 = list_key(p, a * p.C + lane + 64).x;
        RowDir dir;
        dir.load(p.hdr + a * ALLS_DIR);
        const int32_t n = min(dir.d[0], ALLS_DIR * p.C);
        longest = max(longest, n);
        // A table of at least twice the row's entries (power of two, <= ts).
        int tsr = 128;
        while (tsr < 2 * n && tsr < ts) tsr <<= 1;
        if (defer && tsr < 2 * n) {
            if (lane == 0) defer[atomicAdd(&p.misc[MISC_BIG_COUNT], 1)] = (int32_t)a;
            continue;
        }
        const int shift = 32 - (31 - __clz(tsr));
        for (int i = lane; i < tsr; i += 64) table[i] = -1;
        __syncthreads();
        int32_t mine = 0;
        for (int i = lane; i < n; i += 64) {
            const int32_t col = i < p.C && i < lane + 128 ? (i == lane ? hk0 : hk1) : list_key(p, dir.entry(p, a, i)).x;
            uint32_t slot = col_hash(col, shift);
            for (int probe = 0; probe < tsr; ++probe) {
                const int32_t old = atomicCAS(&table[slot], -1, col);
                if (old == -1) {
                    ++mine;
                    break;
                }
                if (old == col) break;
                slot = (slot + 1) & (uint32_t)(tsr - 1);
            }
        }
        int32_t total;
        wave_excl_scan(mine, &total);
        if (lane == 0) u[a] = total;
        __syncthreads();
    }
    // Reading first keeps the shared counter from serialising every block.
    if (lane == 0 && longest > ld_agent(&p.misc[MISC_MAX_BIG])) atomicMax(&p.misc[MISC_MAX_BIG], longest);
}

// Unique-column count when every row is long (3D lists), rows of up to
// 64 * K entries (the longer ones are listed in `defer`, count in
// MISC_BIG_COUNT, for k_rows_hash_count with a full table).  One wave per
// row, software-pipelined over its rows: while row a is hashed, the keys of
// row a + g (g = grid size) and the header of row a + 2g are in flight, so
// the memory round trip overlaps the LDS work.  Headers are spread over
// lanes (lane l holds word l % ALLS_DIR: a uniform load would be moved to a
// scalar register and waited for at once), every load is issued
// unconditionally (lanes past a row read its slot 0) and every lane stores
// the row's count, so that the number of memory operations between a load
// and its use is fixed and the compiler waits for exactly what is used.
template <int K, bool SG>
__global__ __launch_bounds__(64) void k_rows_count_long(ListPool p, int32_t* u, int32_t* defer) {
    constexpr int CAP = 64 * K;
    constexpr int TS = 2 * CAP;
    static_assert((TS & (TS - 1)) == 0 && TS % 256 == 0, "table size: a power of two, 4 ints per lane per step");
    __shared__ __align__(16) int32_t table[TS];
    const int lane = threadIdx.x;
    constexpr int shift = 32 - (31 - __builtin_clz(TS));
    const int64_t g = gridDim.x;
    const int64_t last = p.n_nodes - 1;
    const int hl = lane & (ALLS_DIR - 1);
    int32_t longest = 0;
    int64_t a = xcd_block(blockIdx.x, gridDim.x);
    int32_t ha = p.hdr[min(a, last) * ALLS_DIR + hl];       // header of row a
    int32_t hb = p.hdr[min(a + g, last) * ALLS_DIR + hl];   // header of row a + g
    // Segment lists: each lane's strip element ids of a row's home chunk
    // (no header needed), then the columns from the connectivity.  Entry i
    // of the home chunk is column i % 20 of strip i / 20; lanes past the
    // chunk read strip 0.
    int32_t seg[K], sj[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = lane + 64 * k;
        seg[k] = SG && i < p.C ? i / SEG20 : 0;
        sj[k] = SG && i < p.C ? i - SEG20 * (i / SEG20) : 0;
    }
    int32_t gb[K];   // SG: strip element ids of row a + g
    int32_t kc[K];
    {
        const int lim = min(min(__builtin_amdgcn_readlane(ha, 0), ALLS_DIR * p.C), p.C);
        if constexpr (SG) {
            int32_t ga[K];
#pragma unroll
            for (int k = 0; k < K; ++k) ga[k] = p.geid[min(a, last) * p.gpc + seg[k]];
#pragma unroll
            for (int k = 0; k < K; ++k) gb[k] = p.geid[min(a + g, last) * p.gpc + seg[k]];
#pragma unroll
            for (int k = 0; k < K; ++k) kc[k] = p.conn[(int64_t)seg_elem(p, ga[k]) * p.npe + sj[k]];
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = lane + 64 * k;
                kc[k] = p.keys[min(a, last) * p.C + (i < lim ? i : 0)].x;
            }
        }
        // Settle the prologue: the loop's waits are then those of its own
        // iterations (the compiler merges the entry and back-edge states).
        __builtin_amdgcn_s_waitcnt(0);
    }
    for (; a < p.n_nodes; a += g) {
        const int32_t n = min(__builtin_amdgcn_readlane(ha, 0), ALLS_DIR * p.C);
        const int32_t hd = ha;
        // Loads for the rows ahead.
        const int32_t hc = p.hdr[min(a + 2 * g, last) * ALLS_DIR + hl];
        int32_t kn[K], gc[K];
        if constexpr (SG) {
            // Strip ids of row a + 2g, columns of row a + g (from its ids,
            // loaded an iteration ago).
#pragma unroll
            for (int k = 0; k < K; ++k) gc[k] = p.geid[min(a + 2 * g, last) * p.gpc + seg[k]];
#pragma unroll
            for (int k = 0; k < K; ++k) kn[k] = p.conn[(int64_t)seg_elem(p, gb[k]) * p.npe + sj[k]];
        } else {
            const int64_t b = min(a + g, last);
            const int lim = min(min(__builtin_amdgcn_readlane(hb, 0), ALLS_DIR * p.C), p.C);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = lane + 64 * k;
                kn[k] = p.keys[b * p.C + (i < lim ? i : 0)].x;
            }
        }
        longest = max(longest, n);
        const bool fits = n <= CAP;
        if (!fits && lane == 0) defer[atomicAdd(&p.misc[MISC_BIG_COUNT], 1)] = (int32_t)a;
        if (fits && n > p.C) {   // overflow chunks
            RowDir dir;
#pragma unroll
            for (int j = 0; j < ALLS_DIR; ++j) dir.d[j] = __builtin_amdgcn_readlane(hd, j);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = lane + 64 * k;
                if (i >= p.C && i < n) kc[k] = list_key(p, dir.entry(p, a, i)).x;
            }
            // Settle this rare path's loads here, so that the common path's
            // waits are not merged with it.
            __builtin_amdgcn_s_waitcnt(0);
        }
#pragma unroll
        for (int j = 0; j < TS / 256; ++j) reinterpret_cast<int4*>(table)[j * 64 + lane] = make_int4(-1, -1, -1, -1);
        __syncthreads();
        int32_t mine = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            if (fits && i < n) {
                uint32_t slot = col_hash(kc[k], shift);
                for (int probe = 0; probe < TS; ++probe) {
                    const int32_t old = atomicCAS(&table[slot], -1, kc[k]);
                    if (old == -1) {
                        ++mine;
                        break;
                    }
                    if (old == kc[k]) break;
                    slot = (slot + 1) & (uint32_t)(TS - 1);
                }
            }
        }
        int32_t total;
        wave_excl_scan(mine, &total);
        u[a] = total;   // a deferred row's count is rewritten by the second launch
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            kc[k] = kn[k];
            if constexpr (SG) gb[k] = gc[k];
        }
        ha = hb;
        hb = hc;
    }
    if (lane == 0 && longest > ld_agent(&p.misc[MISC_MAX_BIG])) atomicMax(&p.misc[MISC_MAX_BIG], longest);
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

// The same row_ptr as k_scan_apply for DPN = 2 or 3 and 16-byte aligned u
// and row_ptr.  A thread's 16 nodes own 16*DPN consecutive row_ptr entries
// (64*DPN or 128*DPN bytes from a 16-byte boundary): they are loaded and
// stored as 16-byte vectors instead of one scattered 4- or 8-byte access per
// entry, which left each store instruction on 64 separate cache lines.  The
// thread whose run crosses n takes the per-entry path.
template <typename RP, int DPN>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply_vec(const int32_t* u, int64_t n, const int64_t* tiles,
                                                                 RP* row_ptr) {
    static_assert(SCAN_PER % 4 == 0, "whole int4 loads");
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    const bool whole = base + SCAN_PER <= n;
    int32_t loc[SCAN_PER];
    if (whole) {
        const int4* u4 = reinterpret_cast<const int4*>(u + base);
#pragma unroll
        for (int q = 0; q < SCAN_PER / 4; ++q) {
            const int4 x = u4[q];
            loc[4 * q] = x.x; loc[4 * q + 1] = x.y; loc[4 * q + 2] = x.z; loc[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_PER; ++i) loc[i] = base + i < n ? u[base + i] : 0;
    }
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) s += loc[i];
    int64_t tot;
    int64_t off = tiles[blockIdx.x] + block_excl_scan<int64_t>(s, &tot);
    if (whole) {
        RP v[SCAN_PER * DPN];
#pragma unroll
        for (int i = 0; i < SCAN_PER; ++i) {
#pragma unroll
            for (int dr = 0; dr < DPN; ++dr) v[i * DPN + dr] = (RP)(DPN * (off * DPN + (int64_t)dr * loc[i]));
            off += loc[i];
        }
        RP* dst = row_ptr + base * DPN;
        if constexpr (sizeof(RP) == 4) {
#pragma unroll
            for (int c = 0; c < SCAN_PER * DPN / 4; ++c)
                reinterpret_cast<int4*>(dst)[c] = make_int4(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
        } else {
#pragma unroll
            for (int c = 0; c < SCAN_PER * DPN / 2; ++c)
                reinterpret_cast<longlong2*>(dst)[c] = make_longlong2(v[2 * c], v[2 * c + 1]);
        }
        if (base + SCAN_PER == n) row_ptr[n * DPN] = (RP)(off * DPN * DPN);
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_PER; ++i) {
            const int64_t a = base + i;
            if (a < n) {
#pragma unroll
                for (int dr = 0; dr < DPN; ++dr) row_ptr[a * DPN + dr] = (RP)(DPN * (off * DPN + (int64_t)dr * loc[i]));
                off += loc[i];
                if (a == n - 1) row_ptr[n * DPN] = (RP)(off * DPN * DPN);
            }
        }
    }
}

template <int B, typename VT, bool WRITE, typename RP, typename OT = double>
void rows_dispatch(const ListPool& p, int G, int32_t max_big, const int32_t* big_list, int32_t* u, int32_t* bl,
                   const RP* rp, int32_t* ci, OT* v, int dpn, hipStream_t s) {
    if (G > 0) {
        // Counting is bound by load latency: two rounds of rows per wave
        // keep more loads in flight (about 8% faster on Q4).  The write pass
        // measured slower that way and takes one round.
        constexpr int R = WRITE ? 1 : 2;
        const int64_t rows_per_wave = (int64_t)(64 / G) * R;
        const int64_t threads = (p.n_nodes + rows_per_wave - 1) / rows_per_wave * 64;
        const unsigned blocks = (unsigned)((threads + 255) / 256);
        if constexpr (B == 4 && !WRITE) {
            if (p.ng == 4 && G == 16) {   // Q4 groups: a quad of lanes per row
                const unsigned qb = (unsigned)((p.n_nodes * 4 + 255) / 256);
                hipLaunchKernelGGL((k_count_quad<B, VT>), dim3(qb), dim3(256), 0, s, p, u, bl);
                goto long_rows;
            }
        }
#define ALLS_SMALL(GG, NGG) \
        hipLaunchKernelGGL((k_rows_small<GG, R, NGG, B, VT, WRITE, RP, OT>), dim3(blocks), dim3(256), 0, s, p, u, bl, rp, ci, v, dpn)
        if constexpr (B == 4) {   // 2D: Q4 groups of 4, T3 line groups, mixed keys
            if (G == 16) {
                if (p.ng == 4) ALLS_SMALL(16, 4);
                else if (p.ng == 3) ALLS_SMALL(16, 3);
                else ALLS_SMALL(16, 0);
            } else if (G == 32) {
                if (p.ng == 4) ALLS_SMALL(32, 4);
                else if (p.ng == 3) ALLS_SMALL(32, 3);
                else ALLS_SMALL(32, 0);
            } else {
                if (p.ng == 3) ALLS_SMALL(64, 3);
                else ALLS_SMALL(64, 0);
            }
        } else {
            if (G == 16) ALLS_SMALL(16, 0);
            else if (G == 32) ALLS_SMALL(32, 0);
            else ALLS_SMALL(64, 0);
        }
#undef ALLS_SMALL
    }
long_rows:
    // Long rows: over the big-row list (G > 0) or over every row (G == 0).
    int64_t grid = G > 0 ? 2048 : p.n_nodes;
    if (grid > 262144) grid = 262144;
    if (grid < 1) grid = 1;
    if (!WRITE) {
        int cap = 1;
        while (cap < ALLS_DIR * p.C) cap <<= 1;
        const int ts = 2 * cap;
        if (G > 0 || ts <= 512) {
            hipLaunchKernelGGL(k_rows_hash_count, dim3((unsigned)grid), dim3(64), (size_t)ts * sizeof(int32_t), s,
                               p, G > 0 ? big_list : nullptr, u, ts, nullptr);
            return;
        }
        // Every row (Hex20): rows of up to 256 entries with a 512-slot table
        // (2 KB of LDS), then the rows beyond it, listed in big_list, with
        // the full table.  The count kernel is pipelined over each wave's
        // rows, so it wants several rows per wave: 65536 waves measured
        // 0.62 ms at 80^3 against 0.67 ms for 262144 and 1.18 ms for one
        // row per wave.
        if (p.ng == SEG20)
            hipLaunchKernelGGL((k_rows_count_long<4, true>), dim3((unsigned)std::min<int64_t>(grid, 65536)), dim3(64),
                               0, s, p, u, bl);
        else
            hipLaunchKernelGGL((k_rows_count_long<4, false>), dim3((unsigned)std::min<int64_t>(grid, 65536)), dim3(64),
                               0, s, p, u, bl);
        hipLaunchKernelGGL(k_rows_hash_count, dim3(2048), dim3(64), (size_t)ts * sizeof(int32_t), s, p, bl, u, ts,
                           nullptr);
        return;
    }
    if (max_big <= 0) return;   // no long rows
    int cap = 64;
    while (cap < max_big) cap <<= 1;
    const int32_t* bl_or_all = G > 0 ? big_list : nullptr;
    if (cap <= 256) {
        // Entries per lane: the fewer, the less LDS and registers per wave
        // (Hex20 rows of at most 160 entries take 3: 4 -> 5 waves/SIMD).
        const int epl = (max_big + 63) / 64;
        if constexpr (B == 9 && WRITE) {
            const int ks = (max_big + 59) / 60;   // register slots of 60 entries
            if (G == 0 && p.npe == 20 && ks >= 2 && ks <= 4) {
                // Hex20 rows: segment merge, then the rows it defers (usually
                // none) by the generic kernel over the deferred list.  One
                // row per wave: at 80^3 a grid of every row measured 5.3 ms
                // (FP64) against 6.1-6.5 ms for 262144 waves looping over
                // 8 rows each, whose next row waits for the last one's stores.
                grid = std::max<int64_t>(p.n_nodes, 1);
#define ALLS_SEG(KS)                                                                                        \
    do {                                                                                                    \
        if (p.ng == SEG20)                                                                                  \
            hipLaunchKernelGGL((k_rows_merge_seg<VT, RP, KS, 20, OT, true>), dim3((unsigned)grid), dim3(64), 0, \
                               s, p, rp, ci, v, bl);                                                        \
        else                                                                                                \
            hipLaunchKernelGGL((k_rows_merge_seg<VT, RP, KS, 20, OT, false>), dim3((unsigned)grid), dim3(64), 0, \
                               s, p, rp, ci, v, bl);                                                        \
    } while (0)
                if (ks == 2)
                    ALLS_SEG(2);
                else if (ks == 3)
                    ALLS_SEG(3);
                else
                    ALLS_SEG(4);
#undef ALLS_SEG
                hipLaunchKernelGGL((k_rows_merge_reg<B, VT, RP, 4, OT>), dim3(1024), dim3(64), 0, s, p, bl, rp, ci, v,
                                   MISC_DEFER);
                return;
            }
        }
        if (epl == 1)
            hipLaunchKernelGGL((k_rows_merge_reg<B, VT, RP, 1, OT>), dim3((unsigned)grid), dim3(64), 0, s, p, bl_or_all,
                               rp, ci, v, MISC_BIG_COUNT);
        else if (epl == 2)
            hipLaunchKernelGGL((k_rows_merge_reg<B, VT, RP, 2, OT>), dim3((unsigned)grid), dim3(64), 0, s, p, bl_or_all,
                               rp, ci, v, MISC_BIG_COUNT);
        else if (epl == 3)
            hipLaunchKernelGGL((k_rows_merge_reg<B, VT, RP, 3, OT>), dim3((unsigned)grid), dim3(64), 0, s, p, bl_or_all,
                               rp, ci, v, MISC_BIG_COUNT);
        else
            hipLaunchKernelGGL((k_rows_merge_reg<B, VT, RP, 4, OT>), dim3((unsigned)grid), dim3(64), 0, s, p, bl_or_all,
                               rp, ci, v, MISC_BIG_COUNT);
        return;
    }
    if (cap <= MERGE_CAP_MAX) {
        hipLaunchKernelGGL((k_rows_merge<B, VT, RP, OT>), dim3((unsigned)grid), dim3(64), merge_lds(cap), s, p,
                           bl_or_all, rp, ci, v, cap);
        return;
    }
    // Rows of more than MERGE_CAP_MAX entries (a node shared by very many
    // elements): sort staged in LDS, up to LONG_CAP_MAX entries.
    const size_t lds = (size_t)cap * (sizeof(uint64_t) + sizeof(uint16_t));
    hipLaunchKernelGGL((k_rows_big<B, VT, WRITE, RP, OT>), dim3((unsigned)grid), dim3(64), lds, s, p,
                       G > 0 ? big_list : nullptr, u, rp, ci, v, dpn, cap);
}

}  // namespace

namespace alls {

int32_t long_row_limit() { return LONG_CAP_MAX; }

int64_t scan_tile_count(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

int64_t fused_tile_count(int64_t n_nodes, int G) {
    const int64_t rpt = (int64_t)(256 / G) * fused_iter(G);
    return (n_nodes + rpt - 1) / rpt;
}

void launch_rows_fused(const ListPool& p, int G, int64_t* tstat, void* row_ptr, int i64, int32_t* col_ind,
                       double* val, hipStream_t s) {
    const int64_t nt = fused_tile_count(p.n_nodes, G);
    if (nt <= 0) return;
#define ALLS_FU(GG, NGG, RP)                                                                               \
    hipLaunchKernelGGL((k_rows_fused<GG, NGG, RP>), dim3((unsigned)nt), dim3(256), 0, s, p, tstat, nt,       \
                       static_cast<RP*>(row_ptr), col_ind, val)
#define ALLS_FU_NG(GG, RP)                     \
    do {                                       \
        if (p.ng == 4) ALLS_FU(GG, 4, RP);     \
        else if (p.ng == 3) ALLS_FU(GG, 3, RP); \
        else ALLS_FU(GG, 0, RP);               \
    } while (0)
    if (G == 16) {
        if (i64) ALLS_FU_NG(16, int64_t);
        else ALLS_FU_NG(16, int32_t);
    } else {
        if (i64) ALLS_FU_NG(32, int64_t);
        else ALLS_FU_NG(32, int32_t);
    }
#undef ALLS_FU_NG
#undef ALLS_FU
}

void launch_count(const ListPool& p, int G, int f32vals, int32_t* u, int32_t* big_list, hipStream_t s) {
    if (p.n_nodes <= 0) return;   // no rows (the row kernels clamp row indices to the last row)
    if (p.B == 4)
        rows_dispatch<4, double, false, int32_t, double>(p, G, 0, big_list, u, big_list, nullptr, nullptr, nullptr, 2, s);
    else if (f32vals)
        rows_dispatch<9, float, false, int32_t, double>(p, G, 0, big_list, u, big_list, nullptr, nullptr, nullptr, 3, s);
    else
        rows_dispatch<9, double, false, int32_t, double>(p, G, 0, big_list, u, big_list, nullptr, nullptr, nullptr, 3, s);
}

void launch_scan_total(const int32_t* u, int64_t n, int64_t* tiles, int64_t ntiles, hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tiles, ntiles);
}

void launch_scan_apply(const int32_t* u, int64_t n, const int64_t* tiles, int dpn, void* row_ptr, int i64,
                       hipStream_t s) {
    const int64_t ntiles = scan_tile_count(n);
    if (ntiles == 0) {
        // No nodes: row_ptr = {0} (a fill kernel, so that graphs may hold it).
        launch_fill_i32(static_cast<int32_t*>(row_ptr), i64 ? 2 : 1, 0, s);
        return;
    }
    const bool aligned = ((reinterpret_cast<uintptr_t>(u) | reinterpret_cast<uintptr_t>(row_ptr)) & 15) == 0;
    if (aligned && (dpn == 2 || dpn == 3)) {
        const dim3 g((unsigned)ntiles), b(SCAN_THREADS);
        if (i64 && dpn == 2)
            hipLaunchKernelGGL((k_scan_apply_vec<int64_t, 2>), g, b, 0, s, u, n, tiles, static_cast<int64_t*>(row_ptr));
        else if (i64)
            hipLaunchKernelGGL((k_scan_apply_vec<int64_t, 3>), g, b, 0, s, u, n, tiles, static_cast<int64_t*>(row_ptr));
        else if (dpn == 2)
            hipLaunchKernelGGL((k_scan_apply_vec<int32_t, 2>), g, b, 0, s, u, n, tiles, static_cast<int32_t*>(row_ptr));
        else
            hipLaunchKernelGGL((k_scan_apply_vec<int32_t, 3>), g, b, 0, s, u, n, tiles, static_cast<int32_t*>(row_ptr));
        return;
    }
    if (i64)
        hipLaunchKernelGGL(k_scan_apply<int64_t>, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles, dpn,
                           static_cast<int64_t*>(row_ptr));
    else
        hipLaunchKernelGGL(k_scan_apply<int32_t>, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, u, n, tiles, dpn,
                           static_cast<int32_t*>(row_ptr));
}

void launch_write(const ListPool& p, int G, int f32vals, int dpn, int32_t max_big, int32_t* big_list,
                  const void* row_ptr, int i64, int32_t* col_ind, void* val, int val_f32, hipStream_t s) {
    (void)dpn;
    if (p.n_nodes <= 0) return;   // no rows
    const int32_t mb = max_big;
    double* vd = static_cast<double*>(val);
    float* vf = static_cast<float*>(val);
#define ALLS_W3(VT, RP, rp)                                                                                      \
    do {                                                                                                         \
        if (val_f32) rows_dispatch<9, VT, true, RP, float>(p, G, mb, big_list, nullptr, big_list, rp, col_ind, vf, 3, s); \
        else rows_dispatch<9, VT, true, RP, double>(p, G, mb, big_list, nullptr, big_list, rp, col_ind, vd, 3, s);        \
    } while (0)
    if (i64) {
        const int64_t* rp = static_cast<const int64_t*>(row_ptr);
        if (p.B == 4) rows_dispatch<4, double, true, int64_t>(p, G, mb, big_list, nullptr, big_list, rp, col_ind, vd, 2, s);
        else if (f32vals) ALLS_W3(float, int64_t, rp);
        else ALLS_W3(double, int64_t, rp);
    } else {
        const int32_t* rp = static_cast<const int32_t*>(row_ptr);
        if (p.B == 4) rows_dispatch<4, double, true, int32_t>(p, G, mb, big_list, nullptr, big_list, rp, col_ind, vd, 2, s);
        else if (f32vals) ALLS_W3(float, int32_t, rp);
        else ALLS_W3(double, int32_t, rp);
    }
#undef ALLS_W3
}

}  // namespace alls
