// insert_2d.hip -- Q4 / T3 / mixed plane-stress kernels (gfx950):
//   k_insert_2d  : element integration fused with A-LLS list insertion;
//   k_map        : element -> CSR-slot map of a cached pattern (2D; Hex20
//                  patterns take k_map_rows, a node row at a time);
//   k_scatter_2d : coloured (plain "+=") or FP64-atomic scatter through
//                  that map;
//   k_gather_2d  : value-only re-assembly, one thread per node row.
//
// k_insert_2d replaces assembleGlobalStiffnessKernelALLS (Acc:394-468,
// T3:227-290, Mix:248-312).  The reference walks each row's linked list per
// scalar contribution and merges equal columns with FP64 atomics.  Here one
// thread integrates one element (Elem2D) and, for each of its row nodes a,
// appends one entry per column node b carrying the 2x2 block K_ab: one
// reservation atomic per (element, row node) instead of a list walk and an
// atomic per scalar.  Equal columns are merged by the per-row sort in
// consolidation.  k_scatter_2d replaces the coloured kernels
// (Acc:504-572, T3:316-363, Mix:341-390).
#include "elem2d.h"

using alls_dev::Elem2D;
using alls_dev::RowMoments;

namespace {

typedef double f64x2 __attribute__((ext_vector_type(2)));

// One wave's LDS stage of one row slot i: every lane's (up to) 4 entries of
// its element's row node nd[i].
// Rows are padded (9 x 16 B) so that the lane-strided staging writes spread
// over the LDS banks.
struct RowStage {
    double2 v[64][9];      // entry j as (K00, K01), (K10, K11) (T3: [6], [7] hold the
                           // line group's key block); [8] is padding
    int32_t row[64], pos[64], nn[64];
};

// Triangle and mixed meshes merge the wave's reservations by row node
// before any global atomic: an LDS hash set of the row nodes the wave's
// elements name, which lives in the wave's RowStage (filled only later, by
// the store phase).  Every (element, row node) adds its nn slots to its
// node's count with an LDS atomic that returns its offset; one lane per
// distinct node then takes the whole count from the node's list with ONE
// global atomic, and each element's slots are that base plus its offset.
// 64 consecutive triangles name about 66 distinct rows for their 192
// reservations (C3 1500x1500 T3: 13.5M reservations, 4.64M atomics).  The
// order of the offsets within a node is that of the LDS atomics; only the
// entries of one element in one row (an element that lists a node twice)
// need a fixed order, and those come from one lane in program order.
// Consolidation sums each column in element order whatever the slot order,
// so the values do not change.
constexpr int MATCH_SLOTS = 512;   // >= 2x the most keys of a wave (64 x 4)
struct MatchTable {
    int32_t key[MATCH_SLOTS];      // row node, -1 = empty
    int32_t cnt[MATCH_SLOTS];      // slots requested, then the list position of the first
};
static_assert(sizeof(MatchTable) <= sizeof(RowStage), "the match table lives in the row stage");

// Slot of `key` in the table (inserted if new), linear probing.
__device__ __forceinline__ int32_t match_insert(MatchTable& t, int32_t key) {
    uint32_t s = ((uint32_t)key * 0x9E3779B1u) >> (32 - 9);
    for (int probe = 0; probe < MATCH_SLOTS; ++probe) {
        const int32_t old = atomicCAS(&t.key[s], -1, key);
        if (old == -1 || old == key) break;
        s = (s + 1) & (MATCH_SLOTS - 1);
    }
    return (int32_t)s;
}

// MODE 0: all Q4, 1: all T3, 2: mixed (per-element type array).
//
// A lane integrates one element and reserves nn slots in each of its nn
// node rows (phase 1), then publishes the overflow chunks its reservations
// own (phase 2); a chunk is always published by the wave that reserved it,
// before that wave waits on anything, so no wait can hang on a wave that is
// not running.  The values are then written row slot by row slot through
// LDS: 8 lanes store one element's row group (4 entries x 32 B, one 128-B
// segment), instead of every lane scattering its own 16-B pieces over 64
// rows.  Scattered stores kept the address unit busy (rocprof TA_TA_BUSY ~80%
// of the kernel); grouped stores touch about a sixth of the cache lines per
// instruction.  For Q4 meshes the key is the group's element id (4 B per
// element row, stored straight from the element's lane): the column of
// entry j is the element's node j, so the entries carry no key of their own.
// Per-entry keys cost 0.27 ms of 0.85 at Q4 2000x2000 (4M extra cache lines
// touched by partial writes); the element ids touch 0.5M.  T3 meshes write
// each element row as one whole 128-byte line (3 value blocks and a key
// block; ListPool ng == 3): 96-byte groups and separate 24-byte keys split
// into partial lines at random rows, which measured 0.91 ms for the stores
// alone at 1500x1500 with permuted node ids against 0.38 ms as whole lines.
// Mixed meshes (groups of 3 and 4) keep per-entry keys, stored by the
// element's lane.
//
// Grid-stride over blocks of elements (persistent grid, or one thread per
// element): `base` is uniform over the block, so every wave takes the same
// iterations and barriers; lanes past the end only help with the stores.
template <int MODE>
__global__ __launch_bounds__(256) void k_insert_2d(DevMesh m, ListPool p) {
    extern __shared__ __align__(16) unsigned char lds_insert[];
    RowStage& st = reinterpret_cast<RowStage*>(lds_insert)[threadIdx.x >> 6];
    MatchTable& mt = reinterpret_cast<MatchTable&>(st);
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = xcd_block(blockIdx.x, gridDim.x) * blockDim.x; base < m.n_elems; base += stride) {
        const int64_t e = base + threadIdx.x;
        Elem2D el;
        int nn = 0;
        int32_t pos[4] = {0, 0, 0, 0};
        if constexpr (MODE == 0) {
            // Structured quadrilaterals: one reservation per (element, row
            // node), issued before the integration so that their round trips
            // overlap its arithmetic.  Merging them by wave as below measured
            // slower here (Q4 2000x2000 insertion 0.545 -> 0.61 ms, one box):
            // the table work costs more than the atomics it saves, which are
            // not what bounds this kernel.
            if (e < m.n_elems && el.load_geom<MODE>(m, e, p.misc)) {
                nn = el.nn;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (i < nn) pos[i] = list_reserve_issue(p, el.nd[i], nn);
                el.moments();
                el.settle();
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (i < nn) list_note_long(p, pos[i], nn);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (i < nn) list_own_chunks(p, el.nd[i], pos[i], nn);
            }
        } else {
            // Triangles and mixed meshes: reservations merged by row node
            // (MatchTable; C3 1500x1500 T3 insertion 0.97 -> 0.88 ms, mixed
            // 0.84 -> 0.82 ms, same box).
#pragma unroll
            for (int q = 0; q < MATCH_SLOTS / 64; ++q) {
                mt.key[lane + 64 * q] = -1;
                mt.cnt[lane + 64 * q] = 0;
            }
            const bool ok = e < m.n_elems && el.load_geom<MODE>(m, e, p.misc);
            __syncthreads();
            // (table slot << 16 | offset in the node's merged reservation) per row node
            int32_t so[4] = {0, 0, 0, 0};
            if (ok) {
                nn = el.nn;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i < nn) {
                        const int32_t sl = match_insert(mt, el.nd[i]);
                        so[i] = (sl << 16) | atomicAdd(&mt.cnt[sl], nn);
                    }
                }
            }
            __syncthreads();
            // One global reservation per distinct node, issued before the
            // integration, which overlaps its round trip.
            int32_t gpos[MATCH_SLOTS / 64];
#pragma unroll
            for (int q = 0; q < MATCH_SLOTS / 64; ++q) {
                const int32_t k = mt.key[lane + 64 * q];
                gpos[q] = k >= 0 ? list_reserve_issue(p, k, mt.cnt[lane + 64 * q]) : 0;
            }
            if (ok) {
                el.moments();
                el.settle();
            }
#pragma unroll
            for (int q = 0; q < MATCH_SLOTS / 64; ++q) {
                const int s = lane + 64 * q;
                const int32_t k = mt.key[s];
                if (k >= 0) {
                    const int32_t c = mt.cnt[s];
                    list_note_long(p, gpos[q], c);
                    list_own_chunks(p, k, gpos[q], c);
                    mt.cnt[s] = gpos[q];
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (i < nn) pos[i] = mt.cnt[so[i] >> 16] + (so[i] & 0xffff);
            __syncthreads();   // the row stage is refilled below
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        // Triangles have 3 row nodes; the loop over a fourth is skipped.
#pragma unroll
        for (int i = 0; i < (MODE == 1 ? 3 : 4); ++i) {
            const bool has = i < nn;
            st.nn[lane] = has ? nn : 0;
            if (has) {
                st.row[lane] = el.nd[i];
                st.pos[lane] = pos[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j < nn) {
                        double k[4];
                        el.block(i, j, k);
                        st.v[lane][2 * j] = make_double2(k[0], k[1]);
                        st.v[lane][2 * j + 1] = make_double2(k[2], k[3]);
                    }
                }
                if (MODE == 1 && p.ng == 3) {
                    // Key block of the line group: {element, node 0, node 1,
                    // node 2} as raw bits, then 16 unused bytes (zero), so
                    // that the 8 pieces cover the whole 128-byte line.
                    st.v[lane][6] = make_double2(
                        __longlong_as_double((int64_t)(((uint64_t)(uint32_t)el.nd[0] << 32) | (uint32_t)e)),
                        __longlong_as_double((int64_t)(((uint64_t)(uint32_t)el.nd[2] << 32) | (uint32_t)el.nd[1])));
                    st.v[lane][7] = make_double2(0.0, 0.0);
                } else if constexpr (MODE == 0) {
                    list_store_group(p, el.nd[i], pos[i], (int32_t)e);
                } else {
                    list_store_keys(p, el.nd[i], pos[i], nn, el.nd, (int32_t)e);
                }
            }
            __syncthreads();
            if (MODE == 1 && p.ng == 3) {
                // T3 line groups: 8 lanes write one element row's 128-byte
                // line (3 value blocks + key block).  Every reservation is a
                // whole group (3 slots from a multiple of 3), so a group
                // never straddles chunks.
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int piece = t * 64 + lane;
                    const int src = piece >> 3, w = piece & 7;
                    if (st.nn[src] > 0) {
                        const int32_t s = st.pos[src];
                        const int64_t ch = list_chunk(p, st.row[src], s);
                        if (ch >= 0) {
                            const int64_t line = ch * p.gpc + (s - chunk_of(p, s) * p.C) / 3;
                            // Non-temporal, as the other layouts' value
                            // stores: each store instruction covers whole
                            // lines (C3 T3 1500x1500: 0.77 -> 0.45 ms).
                            const double2 x = st.v[src][w];
                            __builtin_nontemporal_store(f64x2{x.x, x.y}, reinterpret_cast<f64x2*>(p.vals) + line * 8 + w);
                        }
                    }
                }
                __syncthreads();
                continue;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {   // values: lane -> (element src, entry j, half)
                const int piece = t * 64 + lane;
                const int src = piece >> 3, w = piece & 7, j = w >> 1;
                if (j < st.nn[src]) {
                    const int32_t s = st.pos[src] + j;
                    const int64_t ch = list_chunk(p, st.row[src], s);
                    if (ch >= 0) {
                        const int64_t idx = ch * p.C + (s - chunk_of(p, s) * p.C);
                        // Non-temporal: the lists are read once, by another
                        // kernel (Q4 2000x2000: 0.58 -> 0.54 ms, and the
                        // count pass after it gained 6%; mixed 1500x1500:
                        // 0.86 -> 0.84 ms).
                        const double2 x = st.v[src][w];
                        __builtin_nontemporal_store(f64x2{x.x, x.y}, reinterpret_cast<f64x2*>(p.vals) + idx * 2 + (w & 1));
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ void add_val(double* v, double x, bool atomic) {
    if (atomic) atomicAdd(v, x);   // global_atomic_add_f64 (-munsafe-fp-atomics)
    else *v += x;
}

template <int MODE, bool ATOMIC>
__global__ __launch_bounds__(256) void k_scatter_2d(DevMesh m, Scatter sc, int32_t* misc) {
    const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= sc.count) return;
    const int64_t e = sc.elist ? (int64_t)sc.elist[t] : t;
    Elem2D el;
    if (!el.load<MODE>(m, e, misc)) return;
    const int32_t* mp = sc.map + e * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= el.nn) continue;
        const int64_t a = el.nd[i];
        const int64_t rs = rp_at(sc.row_ptr, sc.i64, 2 * a);
        const int64_t rowlen = rp_at(sc.row_ptr, sc.i64, 2 * a + 1) - rs;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= el.nn) continue;
            double k[4];
            el.block(i, j, k);
            double* v = sc.val + rs + 2 * (int64_t)mp[i * 4 + j];
            add_val(v, k[0], ATOMIC);
            add_val(v + 1, k[1], ATOMIC);
            add_val(v + rowlen, k[2], ATOMIC);
            add_val(v + rowlen + 1, k[3], ATOMIC);
        }
    }
}

// Value-only re-assembly of a cached 2D pattern without atomics: one thread
// owns node row a, walks the elements incident to a in ascending element id
// (the serial assembly's order, Acc:336), recomputes each element's moments
// and adds its blocks K_ij (i = local index of a) into the row's blocks.
// A row of at most GATHER_SLOTS block columns (every row of a mesh built
// from quadrilateral cells, split or not) accumulates in registers: block
// slot s is updated by a predicated add, and the row is stored once at the
// end.  Longer rows add into the CSR rows a*2 and a*2+1 in memory, which no
// other thread touches.  Both forms sum 0 + v1 + v2 + ... in the same order.
// An element that lists node a twice appears twice in a's incidence list
// (once per conn slot); it is visited once, and that visit adds the blocks
// of every local index of a in ascending order (row node, then column
// node), the serial order.
constexpr int GATHER_SLOTS = 9;
constexpr int GATHER_BLOCK = 128;   // threads per block (LDS: 9 slots x 32 B each)

// Node id i of an element, i a run-time value: a select chain, so that the
// element's arrays keep constant indices and stay in registers (an indexed
// read would place the whole element record in scratch memory).
__device__ __forceinline__ int32_t nd_at(const Elem2D& el, int i) {
    return i == 0 ? el.nd[0] : i == 1 ? el.nd[1] : i == 2 ? el.nd[2] : el.nd[3];
}

// Integrates element e (loaded) for row node a and adds its blocks into the
// row's block slots in LDS (slot from the cached map): one 32-byte
// read-add-write per block, in the element's column order, so coincident
// columns of one element add one after the other.
__device__ __forceinline__ void gather_add_element(Elem2D& el, int64_t e, int64_t a, const Scatter& sc,
                                                   double4* slots) {
    el.material();
    // Only row node i's moments are formed (i is a run-time index).
#pragma unroll 1
    for (int i = 0; i < el.nn; ++i) {
        if (nd_at(el, i) != a) continue;
        RowMoments rm;
        el.row_moments(i, rm);
        const int4 mp = *reinterpret_cast<const int4*>(sc.map + e * 16 + i * 4);
        const int32_t sl[4] = {mp.x, mp.y, mp.z, mp.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= el.nn) continue;
            double kb[4];
            el.row_block(rm, j, kb);
            double4* d = slots + (int64_t)sl[j] * GATHER_BLOCK;
            const double4 o = *d;
            *d = make_double4(o.x + kb[0], o.y + kb[1], o.z + kb[2], o.w + kb[3]);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(GATHER_BLOCK) void k_gather_2d(DevMesh m, Scatter sc, const int64_t* inc_off,
                                                            const int32_t* inc_list, int32_t* misc) {
    // Block slots of the thread's row, slot q at [q][thread]: consecutive
    // threads, consecutive 32-byte blocks.
    __shared__ double4 gslot[GATHER_SLOTS * GATHER_BLOCK];
    const int64_t a = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (a >= m.n_nodes) return;
    const int64_t r0 = rp_at(sc.row_ptr, sc.i64, 2 * a);
    const int64_t r1 = rp_at(sc.row_ptr, sc.i64, 2 * a + 1);
    const int64_t rowlen = r1 - r0;
    const int64_t k0 = inc_off[a], k1 = inc_off[a + 1];
    if (rowlen <= 2 * GATHER_SLOTS) {
        double4* slots = gslot + threadIdx.x;
#pragma unroll
        for (int q = 0; q < GATHER_SLOTS; ++q) slots[q * GATHER_BLOCK] = make_double4(0.0, 0.0, 0.0, 0.0);
        // Four incident elements at a time (every incident element of an
        // interior quadrilateral row): their ids, node lists and coordinates
        // are loaded together (three dependent rounds of loads per batch
        // instead of per element), then integrated one after the other in
        // ascending id.  The list is sorted, so a repeat is the same element
        // again and is skipped.
        int64_t prev = -1;
        for (int64_t k = k0; k < k1; k += 4) {
            int64_t e[4];
            bool u[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) e[b] = k + b < k1 ? (int64_t)inc_list[k + b] : -1;
            u[0] = e[0] != prev;
#pragma unroll
            for (int b = 1; b < 4; ++b) u[b] = e[b] >= 0 && e[b] != e[b - 1];
#pragma unroll
            for (int b = 0; b < 4; ++b) prev = e[b] >= 0 ? e[b] : prev;
            Elem2D el0, el1, el2, el3;
            u[0] = u[0] && el0.load_nodes<MODE>(m, e[0], misc);
            u[1] = u[1] && el1.load_nodes<MODE>(m, e[1], misc);
            u[2] = u[2] && el2.load_nodes<MODE>(m, e[2], misc);
            u[3] = u[3] && el3.load_nodes<MODE>(m, e[3], misc);
            if (u[0]) el0.load_coords(m, e[0]);
            if (u[1]) el1.load_coords(m, e[1]);
            if (u[2]) el2.load_coords(m, e[2]);
            if (u[3]) el3.load_coords(m, e[3]);
            if (u[0]) gather_add_element(el0, e[0], a, sc, slots);
            if (u[1]) gather_add_element(el1, e[1], a, sc, slots);
            if (u[2]) gather_add_element(el2, e[2], a, sc, slots);
            if (u[3]) gather_add_element(el3, e[3], a, sc, slots);
        }
#pragma unroll
        for (int q = 0; q < GATHER_SLOTS; ++q) {
            if (2 * q < rowlen) {
                const double4 v = slots[q * GATHER_BLOCK];
                *reinterpret_cast<double2*>(sc.val + r0 + 2 * q) = make_double2(v.x, v.y);
                *reinterpret_cast<double2*>(sc.val + r1 + 2 * q) = make_double2(v.z, v.w);
            }
        }
        return;
    }
    const int64_t r2 = rp_at(sc.row_ptr, sc.i64, 2 * a + 2);
    for (int64_t q = r0; q < r2; ++q) sc.val[q] = 0.0;
    int64_t prev = -1;
    for (int64_t k = k0; k < k1; ++k) {
        const int64_t e = inc_list[k];
        if (e == prev) continue;
        prev = e;
        Elem2D el;
        if (!el.load_geom<MODE>(m, e, misc)) continue;
        el.material();
#pragma unroll 1
        for (int i = 0; i < el.nn; ++i) {
            if (nd_at(el, i) != a) continue;
            RowMoments rm;
            el.row_moments(i, rm);
            const int32_t* mp = sc.map + e * 16 + i * 4;
            bool twice = false;   // a column node listed twice by the element
#pragma unroll
            for (int j = 1; j < 4; ++j)
#pragma unroll
                for (int j2 = 0; j2 < j; ++j2) twice |= j < el.nn && mp[j] == mp[j2];
            if (twice) {
                // Coincident columns: load-add-store one block after the
                // other, so the second add sees the first.
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j >= el.nn) continue;
                    double kb[4];
                    el.row_block(rm, j, kb);
                    double* v = sc.val + r0 + 2 * (int64_t)mp[j];
                    const double2 lo = *reinterpret_cast<const double2*>(v);
                    const double2 hi = *reinterpret_cast<const double2*>(v + rowlen);
                    *reinterpret_cast<double2*>(v) = make_double2(lo.x + kb[0], lo.y + kb[1]);
                    *reinterpret_cast<double2*>(v + rowlen) = make_double2(hi.x + kb[2], hi.y + kb[3]);
                }
                continue;
            }
            // All loads of the element's blocks first, then all stores: the
            // element's columns are distinct, and issuing the loads together
            // keeps them in flight instead of chaining load-add-store.
            double* v[4];
            double old[4][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j >= el.nn) continue;
                v[j] = sc.val + r0 + 2 * (int64_t)mp[j];
                const double2 lo = *reinterpret_cast<const double2*>(v[j]);
                const double2 hi = *reinterpret_cast<const double2*>(v[j] + rowlen);
                old[j][0] = lo.x; old[j][1] = lo.y; old[j][2] = hi.x; old[j][3] = hi.y;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j >= el.nn) continue;
                double kb[4];
                el.row_block(rm, j, kb);
                *reinterpret_cast<double2*>(v[j]) = make_double2(old[j][0] + kb[0], old[j][1] + kb[1]);
                *reinterpret_cast<double2*>(v[j] + rowlen) = make_double2(old[j][2] + kb[2], old[j][3] + kb[3]);
            }
        }
    }
}

// 2D: rank of every (row node i, column node j) pair of element e in the
// cached pattern, a thread per (element, row node), by binary search of
// column b*dpn in DOF row a*dpn (the search of the reference's
// binary_search, Acc:214-222, done once per pattern).
template <int NN, typename RP>
__global__ __launch_bounds__(256) void k_map(DevMesh m, const RP* rp, const int32_t* ci, int32_t* map,
                                             int32_t* misc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = t / NN;
    const int i = (int)(t % NN);
    if (e >= m.n_elems) return;
    const int kind = m.types ? (int)m.types[e] : m.uniform_type;
    const int nn = kind == 0 ? 4 : kind == 1 ? 3 : 20;
    if (i >= nn) return;
    const int dpn = m.dim;
    const int32_t* c = m.conn + e * m.npe;
    const int64_t a = c[i];
    const int64_t rs = (int64_t)rp[a * dpn];
    const int32_t u = (int32_t)(((int64_t)rp[a * dpn + 1] - rs) / dpn);   // column nodes of row a
    // The NN binary searches advance together, one probe each per round, so
    // their loads are in flight at the same time instead of one after another.
    int32_t tgt[NN], lo[NN], hi[NN], res[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        tgt[j] = j < nn ? c[j] * dpn : -1;
        lo[j] = 0;
        hi[j] = j < nn ? u - 1 : -1;
        res[j] = -1;
    }
    for (int round = 0; round < 32; ++round) {
        bool open = false;
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            if (res[j] < 0 && lo[j] <= hi[j]) {
                const int32_t mid = (lo[j] + hi[j]) >> 1;
                const int32_t v = ci[rs + (int64_t)mid * dpn];
                if (v == tgt[j]) res[j] = mid;
                else if (v < tgt[j]) lo[j] = mid + 1;
                else hi[j] = mid - 1;
                open |= res[j] < 0 && lo[j] <= hi[j];
            }
        }
        if (!open) break;
    }
    if (i == 0) {   // an element that lists one node twice (the value-only Hex20 gather needs to know)
        bool twice = false;
#pragma unroll
        for (int j = 1; j < NN; ++j)
#pragma unroll
            for (int k = 0; k < j; ++k) twice |= j < nn && tgt[j] == tgt[k];
        if (twice) st_agent(&misc[MISC_ERR + ERR_REPEATED_NODE], 1);
    }
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        if (j >= nn) continue;
        if (res[j] < 0) {
            st_agent(&misc[MISC_ERR + ERR_INVALID_MESH], 1);
            res[j] = 0;
        }
        map[e * NN * NN + i * NN + j] = res[j];
    }
}

// Hex20: the same ranks a node row at a time: G lanes take row a's
// incident elements (inc_list: sorted by id, an element listed once per slot
// that names a), NN lanes per element, lane j searching column node
// conn[e][j] among the row's sorted columns (DOF row a * dpn of col_ind, a
// few cache lines all lanes share).  The element's row slot i is found by a
// ballot of its lanes for conn[e][j] == a; the r-th listing of an element
// that names a twice takes the r-th such slot.  Writes map[e][i][j]; flags an
// element that lists a node twice, and a column missing from the pattern.
template <int NN, int G, typename RP>
__global__ __launch_bounds__(256) void k_map_rows(DevMesh m, const RP* rp, const int32_t* ci, const int64_t* off,
                                                  const int32_t* list, int32_t* map, int32_t* misc) {
    constexpr int SPP = G / NN;    // elements per pass
    constexpr int LPP = SPP * NN;  // lanes per pass (Hex20: 60 of 64)
    static_assert(64 % G == 0 && SPP >= 1, "groups inside a wave, whole elements per pass");
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a = gid / G;
    const int l = (int)(gid % G);
    const int wl = threadIdx.x & 63;
    bool twice = false, missing = false;
    if (a < m.n_nodes) {
        const int dpn = m.dim;
        const int64_t rs = (int64_t)rp[a * dpn];
        const int32_t u = (int32_t)(((int64_t)rp[a * dpn + 1] - rs) / dpn);   // column nodes of row a
        const int64_t lo = off[a], ne = off[a + 1] - lo;
        const int sj = l % NN, sb = wl - sj;   // slot j of this lane, first wave lane of its element
        for (int64_t s0 = 0; s0 < ne; s0 += SPP) {
            const int64_t k = s0 + l / NN;
            const bool act = l < LPP && k < ne;
            int32_t e = 0, nn = 0, cj = -1;
            if (act) {
                e = list[lo + k];
                const int kind = m.types ? (int)m.types[e] : m.uniform_type;
                nn = kind == 0 ? 4 : kind == 1 ? 3 : 20;
                if (sj < nn) cj = m.conn[(int64_t)e * m.npe + sj];
            }
            // The element's slots naming a, as bits relative to its first lane.
            const uint64_t hit = __ballot(cj == (int32_t)a);
            if (!act || sj >= nn) continue;
            uint32_t mine = (uint32_t)(hit >> sb) & ((1u << NN) - 1u);
            int r = 0;
            for (int64_t q = k - 1; q >= 0 && list[lo + q] == e; --q) ++r;   // earlier listings of e in this row
            twice |= r > 0;
            for (int t = 0; t < r; ++t) mine &= mine - 1u;
            if (mine == 0u) {   // the incidence list and the connectivity disagree: never
                missing = true;
                continue;
            }
            const int i = __builtin_ctz(mine);
            const int32_t tgt = cj * dpn;
            int32_t lo2 = 0, hi2 = u - 1, res = -1;
            while (lo2 <= hi2) {
                const int32_t mid = (lo2 + hi2) >> 1;
                const int32_t v = ci[rs + (int64_t)mid * dpn];
                if (v == tgt) {
                    res = mid;
                    break;
                }
                if (v < tgt) lo2 = mid + 1;
                else hi2 = mid - 1;
            }
            if (res < 0) {
                missing = true;
                res = 0;
            }
            map[(int64_t)e * NN * NN + i * NN + sj] = res;
        }
    }
    if (twice) st_agent(&misc[MISC_ERR + ERR_REPEATED_NODE], 1);
    if (missing) st_agent(&misc[MISC_ERR + ERR_INVALID_MESH], 1);
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += stride) p[k] = v;
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = n / 2;
    double2* q = reinterpret_cast<double2*>(p);
    for (int64_t k = i; k < n2; k += stride) q[k] = make_double2(v, v);
    if (i == 0 && (n & 1)) p[n - 1] = v;
}

unsigned grid_for(int64_t n, int block) {
    int64_t b = (n + block - 1) / block;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

namespace alls {

void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_fill_f64(double* p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n / 2 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_insert_2d(const DevMesh& m, const ListPool& p, int mode, int block, int64_t max_blocks, hipStream_t s) {
    if (m.n_elems <= 0) return;
    if (block != 64 && block != 128 && block != 256) block = 256;
    unsigned g = grid_for(m.n_elems, block);
    if (max_blocks > 0 && g > max_blocks) g = (unsigned)max_blocks;
    const size_t lds = (size_t)(block / 64) * sizeof(RowStage);
    if (mode == 0) hipLaunchKernelGGL(k_insert_2d<0>, dim3(g), dim3(block), lds, s, m, p);
    else if (mode == 1) hipLaunchKernelGGL(k_insert_2d<1>, dim3(g), dim3(block), lds, s, m, p);
    else hipLaunchKernelGGL(k_insert_2d<2>, dim3(g), dim3(block), lds, s, m, p);
}

void launch_build_map(const DevMesh& m, int kind, const void* row_ptr, int i64, const int32_t* col_ind,
                      const int64_t* inc_off, const int32_t* inc_list, int32_t* map, int32_t* misc, hipStream_t s) {
    if (m.n_elems <= 0 || m.n_nodes <= 0) return;
    if (kind == 3) {
        // Hex20: a wave per node row, three elements of 20 lanes per pass.
        // 80^3 pattern build 18.0-18.5 -> 16.1-16.2 ms against a thread per
        // (element, row node).  2D keeps the thread per (element, row node):
        // row at a time was the same on Q4 2000^2 and 0.2 ms slower on the
        // permuted-id T3 and mixed 1500^2 meshes (scattered elements per row).
        const unsigned g = grid_for(m.n_nodes * 64, 256);
        if (i64)
            hipLaunchKernelGGL((k_map_rows<20, 64, int64_t>), dim3(g), dim3(256), 0, s, m,
                               static_cast<const int64_t*>(row_ptr), col_ind, inc_off, inc_list, map, misc);
        else
            hipLaunchKernelGGL((k_map_rows<20, 64, int32_t>), dim3(g), dim3(256), 0, s, m,
                               static_cast<const int32_t*>(row_ptr), col_ind, inc_off, inc_list, map, misc);
        return;
    }
    const unsigned g = grid_for(m.n_elems * 4, 256);
    if (i64)
        hipLaunchKernelGGL((k_map<4, int64_t>), dim3(g), dim3(256), 0, s, m, static_cast<const int64_t*>(row_ptr),
                           col_ind, map, misc);
    else
        hipLaunchKernelGGL((k_map<4, int32_t>), dim3(g), dim3(256), 0, s, m, static_cast<const int32_t*>(row_ptr),
                           col_ind, map, misc);
}

void launch_scatter_2d(const DevMesh& m, int kind, const Scatter& sc, int atomic, int32_t* misc, hipStream_t s) {
    if (sc.count <= 0) return;
    const unsigned g = grid_for(sc.count, 256);
#define ALLS_SC(MODE)                                                                              \
    do {                                                                                           \
        if (atomic) hipLaunchKernelGGL((k_scatter_2d<MODE, true>), dim3(g), dim3(256), 0, s, m, sc, misc);  \
        else hipLaunchKernelGGL((k_scatter_2d<MODE, false>), dim3(g), dim3(256), 0, s, m, sc, misc);        \
    } while (0)
    if (kind == 0) ALLS_SC(0);
    else if (kind == 1) ALLS_SC(1);
    else ALLS_SC(2);
#undef ALLS_SC
}

void launch_gather_2d(const DevMesh& m, int kind, const Scatter& sc, const int64_t* inc_off, const int32_t* inc_list,
                      int32_t* misc, hipStream_t s) {
    if (m.n_nodes <= 0) return;
    const unsigned g = grid_for(m.n_nodes, GATHER_BLOCK);
    const dim3 b(GATHER_BLOCK);
    if (kind == 0) hipLaunchKernelGGL(k_gather_2d<0>, dim3(g), b, 0, s, m, sc, inc_off, inc_list, misc);
    else if (kind == 1) hipLaunchKernelGGL(k_gather_2d<1>, dim3(g), b, 0, s, m, sc, inc_off, inc_list, misc);
    else hipLaunchKernelGGL(k_gather_2d<2>, dim3(g), b, 0, s, m, sc, inc_off, inc_list, misc);
}

}  // namespace alls
