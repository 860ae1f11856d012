// alls_internal.h -- shared definitions of liballs (gfx950 HIP + host C++).
//
// Data layout in HBM (one assembly):
//   hdr   int32 [n_nodes][8]   hdr[a][0] = entries appended to node-row a,
//                              hdr[a][k] = 1 + index of overflow chunk k (0 = none), k = 1..7
//   vals  VT    [chunks][C][B] B = dpn*dpn block of the entry (VT = double, or
//                              float for FP32-integrated Hex20)
//   Q4:    geid  int32 [chunks][C/ng]  element id of each group of ng slots
//   T3, mixed 2D and Hex20:
//          keys  int2  [chunks][C]     (column node b, element id e) of every entry
// Chunk a (a < n_nodes) is node-row a's home chunk; overflow chunks are taken
// from a bump counter above n_nodes.  Rows are node rows: one entry carries the
// dpn x dpn block that an element adds to DOF rows a*dpn.. and DOF columns b*dpn..
//
// In a Q4 mesh an element reserves ng = 4 slots in each of its node rows,
// and entry q of the group is its block with column node conn[e][q]: the
// group's slots need no key of their own, only the group's element id.
// Chunks hold whole groups (C is a multiple of 4).  Keys for the rows that
// the wave-per-row kernels take are written out by the count pass (keys[] is
// kept for that).  T3, mixed and Hex20 entries keep per-entry keys (T3
// groups of 3 measured slower: their readers load conn[e] once per group,
// scattered when node ids are permuted; a mixed mesh's groups would differ
// in size, and padding its triangles to 4 slots measured 25% slower).  The
// group code takes ng = 3 too.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define ALLS_DIR 8             // chunks addressable per row (home + 7 overflow)
#define ALLS_SPIN_MAX (1 << 22)

// err[] slots written by kernels, read back by the host after insertion.
enum {
    ERR_DIR_OVERFLOW = 0,      // max entries requested by a row beyond ALLS_DIR*C
    ERR_POOL_EXHAUSTED = 1,
    ERR_INVALID_MESH = 2,
    ERR_SPIN_TIMEOUT = 3,
    ERR_ROW_TOO_BIG = 4,
    ERR_LOOKBACK = 5,          // fused pass: a look-back wait hit its spin limit
    ERR_REPEATED_NODE = 6,     // scatter map: an element lists one node twice (not an error)
    ERR_SLOTS = 8
};
// misc[] layout (int32 each, one buffer so a single copy brings them back).
enum {
    MISC_ERR = 0,              // ERR_SLOTS entries
    MISC_OVF_TOP = 8,          // overflow chunks taken
    MISC_BIG_COUNT = 9,        // rows routed to the wave-per-row kernel
    MISC_MAX_BIG = 10,         // most entries in one of those rows
    MISC_MAX_ROW = 11,         // longest row seen by the insertion, when above ListPool::long_row
    MISC_TICKET = 12,          // tile tickets of the fused small-row pass
    MISC_TOTAL_LO = 13,        // unique node pairs found by the fused pass (low 31 bits)
    MISC_TOTAL_HI = 14,        //   and the bits above
    MISC_DEFER = 15,           // rows the segment merge leaves to k_rows_merge_reg
    MISC_SIZE = 16
};

struct DevMesh {
    const double* coords;
    const int32_t* conn;
    const uint8_t* types;
    const double* E_elem;
    int64_t n_nodes;
    int64_t n_elems;
    int32_t dim;
    int32_t npe;
    int32_t uniform_type;
    int32_t _pad;
    double E, nu, thickness;
};

struct ListPool {
    int32_t* hdr;
    int2* keys;
    void* vals;
    int32_t* misc;
    int32_t* geid;             // 2D: element id of each slot group
    const int32_t* conn;       // 2D: the mesh's connectivity (column nodes of a group)
    int64_t n_nodes;           // home chunks
    int64_t cap_chunks;        // home + overflow capacity
    int64_t n_elems;           // bound of a group's or segment's element id
    int32_t C;                 // entries per chunk (<= ALLS_MAX_CHUNK)
    int32_t B;                 // values per entry
    uint32_t cmagic;           // chunk_magic(C): s / C as a multiply-high
    int32_t long_row;          // reservations ending past this many entries record MISC_MAX_ROW
    int32_t ng;                // 4: Q4 groups (element id per group in geid); 3: T3 line groups
                               // (below); 20: Hex20 segments (below); 0: per-entry keys
    int32_t gpc;               // 2D: groups per chunk (C / ng)
    int32_t npe;               // 2D: conn stride
    int32_t _pad;
};

#define ALLS_MAX_CHUNK 16384

// T3 line groups (ng == 3).  Every reservation of a triangle mesh is one
// group of 3 slots (the element's row block: columns node 0, 1, 2), and a
// chunk holds C = 3 * gpc slots.  Each group lives in ONE 128-byte line of
// `vals`: the three 32-byte value blocks of its slots, then a key block of
// int32 {element, column 0, column 1, column 2} and 16 unused bytes.  The
// insertion writes whole lines (one per element row, no separate key
// array); slot indices (ch * C + r) are the same as for the other layouts,
// and list_val_slot / list_key map them to the line.
constexpr int LINE3_SLOTS = 4;   // value-block slots per 128-byte line group

// Hex20 segments (ng == 20).  Every reservation of a Hex20 mesh is 20 slots
// per row node an element names (one strip per naming), each strip holds the
// element's 20 column blocks in its node order, and chunks hold whole strips
// (C a multiple of 20).  So slot idx = ch * C + r is column conn[e][idx % 20]
// of element e = geid[idx / 20], and the lists carry one 4-byte element id
// per strip instead of an 8-byte key per entry.  An id outside [0, n_elems)
// (a strip of a failed call) reads as element 0, so that no read leaves the
// connectivity.
constexpr int SEG20 = 20;

// Multiplier for s / C without a divide: floor(s * m / 2^32) with
// m = ceil(2^32 / C) equals floor(s / C) for s < 2^32 / C (>= 262144 slots
// for C <= ALLS_MAX_CHUNK) and never falls below it for larger s, so a slot
// past the 8-chunk directory is always seen as past it.
inline uint32_t chunk_magic(int32_t C) { return 0xffffffffu / (uint32_t)C + 1u; }

// Element -> CSR-slot scatter of a cached pattern.  map[e*npe2 + i*npe + j]
// is the rank of column node conn[e][j] among the sorted unique column nodes
// of node row conn[e][i]; DOF entry (dr, dc) of that block sits at
// row_ptr[a*dpn] + rank*dpn + dr*rowlen + dc with rowlen = row_ptr[a*dpn+1] -
// row_ptr[a*dpn].
struct Scatter {
    const int32_t* map;
    const void* row_ptr;
    double* val;
    const int32_t* elist;      // elements to process, or null => 0..count-1
    int64_t count;
    int32_t i64;               // row_ptr is int64
    int32_t npe;               // map stride per node (4 in 2D, 20 for Hex20)
    void* stage;               // Hex20 value-only path: [n_elems*400][9] blocks in element order (T)
};

#ifdef __HIPCC__
#define ALLS_DEV __device__ __forceinline__

// Element id of a Hex20 strip, read as element 0 when out of range.
ALLS_DEV int32_t seg_elem(const ListPool& p, int32_t e) {
    return (uint32_t)e < (uint64_t)p.n_elems ? e : 0;
}

ALLS_DEV int64_t rp_at(const void* rp, int i64, int64_t i) {
    return i64 ? static_cast<const int64_t*>(rp)[i] : (int64_t) static_cast<const int32_t*>(rp)[i];
}

ALLS_DEV int32_t ld_agent(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ALLS_DEV void st_agent(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ALLS_DEV int32_t add_agent(int32_t* p, int32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ALLS_DEV void max_agent(int32_t* p, int32_t v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap so that each
// XCD works on one contiguous 1/8 of the index space: neighbouring elements
// (which share rows) then meet in one XCD's L2.
ALLS_DEV int64_t xcd_block(int64_t bid, int64_t nb) {
    int64_t q = nb >> 3, r = nb & 7, x = bid & 7, i = bid >> 3;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Grid-stride loop over [0, count) that keeps xcd_block's property: block
// bid (XCD bid & 7, the (bid >> 3)-th block there) visits first, first +
// step, ... below end, all inside its XCD's contiguous 1/8 of the range, so
// a persistent grid's XCDs each sweep one slab of elements as a grid of one
// block per element does.  Grids of fewer than 8 blocks stride plainly.
struct XcdRange {
    int64_t first, end, step;
};
ALLS_DEV XcdRange xcd_range(int64_t bid, int64_t nb, int64_t count) {
    if (nb < 8) return XcdRange{bid, count, nb};
    const int64_t x = bid & 7, i = bid >> 3;
    const int64_t nbx = (nb >> 3) + (x < (nb & 7) ? 1 : 0);
    const int64_t cq = count >> 3, cr = count & 7;
    const int64_t lo = x * cq + (x < cr ? x : cr);
    return XcdRange{lo + i, lo + cq + (x < cr ? 1 : 0), nbx};
}

// Reserve n consecutive slots of node-row a; returns the first slot.
ALLS_DEV int32_t chunk_of(const ListPool& p, int32_t s) { return (int32_t)__umulhi((uint32_t)s, p.cmagic); }

ALLS_DEV int32_t list_reserve(const ListPool& p, int32_t a, int32_t n) {
    const int32_t pos = add_agent(&p.hdr[(int64_t)a * ALLS_DIR], n);
    if (pos + n > p.long_row) max_agent(&p.misc[MISC_MAX_ROW], pos + n);
    return pos;
}

// list_reserve in two halves, for a thread that reserves in several rows:
// the reservation atomics go out back to back (no branch on a result
// between them), and the long-row note is taken once all have returned.
ALLS_DEV int32_t list_reserve_issue(const ListPool& p, int32_t a, int32_t n) {
    return add_agent(&p.hdr[(int64_t)a * ALLS_DIR], n);
}
ALLS_DEV void list_note_long(const ListPool& p, int32_t pos, int32_t n) {
    if (pos + n > p.long_row) max_agent(&p.misc[MISC_MAX_ROW], pos + n);
}

// The reservation [pos, pos+n) owns every chunk k >= 1 whose first slot
// k*C lies inside it: take that chunk from the pool and publish it.
ALLS_DEV void list_own_chunks(const ListPool& p, int32_t a, int32_t pos, int32_t n) {
    int32_t k = chunk_of(p, pos + p.C - 1);
    if (k < 1) k = 1;
    for (; k * p.C < pos + n; ++k) {
        if (k >= ALLS_DIR) {
            max_agent(&p.misc[MISC_ERR + ERR_DIR_OVERFLOW], pos + n);
            return;
        }
        int32_t c = add_agent(&p.misc[MISC_OVF_TOP], 1);
        int32_t tag;
        if (p.n_nodes + c >= p.cap_chunks) {
            st_agent(&p.misc[MISC_ERR + ERR_POOL_EXHAUSTED], 1);
            tag = -1;                              // waiters give up on it
        } else {
            tag = c + 1;
        }
        st_agent(&p.hdr[(int64_t)a * ALLS_DIR + k], tag);
    }
}

// Chunk index holding slot s of row a, or -1.  Waits (bounded) for the
// owner of an overflow chunk to publish it.
ALLS_DEV int64_t list_chunk(const ListPool& p, int32_t a, int32_t s) {
    int32_t k = chunk_of(p, s);
    if (k == 0) return a;
    if (k >= ALLS_DIR) return -1;
    const int32_t* d = &p.hdr[(int64_t)a * ALLS_DIR + k];
    int32_t v = ld_agent(d);
    for (int32_t it = 0; v == 0; ++it) {
        if (it >= ALLS_SPIN_MAX) {
            st_agent(&p.misc[MISC_ERR + ERR_SPIN_TIMEOUT], 1);
            return -1;
        }
        __builtin_amdgcn_s_sleep(2);
        v = ld_agent(d);
    }
    if (v < 0) return -1;
    return p.n_nodes + (int64_t)(v - 1);
}

// Keyed lists (mixed 2D meshes): the keys (column node, element) of one
// element's nn entries in row a, slots [pos, pos+nn), stored by the thread
// that reserved them, straight from its registers.
ALLS_DEV void list_store_keys(const ListPool& p, int32_t a, int32_t pos, int nn, const int32_t* nd, int32_t e) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= nn) continue;
        const int32_t s = pos + j;
        const int64_t ch = list_chunk(p, a, s);
        if (ch >= 0) p.keys[ch * p.C + (s - chunk_of(p, s) * p.C)] = make_int2(nd[j], e);
    }
}

// Value-block index of slot idx (ch * C + r): its place in the line groups
// for T3 lists, idx itself otherwise.
ALLS_DEV int64_t list_val_slot(const ListPool& p, int64_t idx) {
    if (p.ng == 3) {
        const int64_t g = idx / 3;
        return g * LINE3_SLOTS + (idx - g * 3);
    }
    return idx;
}
// (column, element) key of slot idx.
ALLS_DEV int2 list_key(const ListPool& p, int64_t idx) {
    if (p.ng == SEG20) {
        const int64_t g = idx / SEG20;
        const int32_t e = seg_elem(p, p.geid[g]);
        return make_int2(p.conn[(int64_t)e * p.npe + (int)(idx - g * SEG20)], e);
    }
    if (p.ng == 3) {
        const int64_t g = idx / 3;
        const int32_t* kb = reinterpret_cast<const int32_t*>(static_cast<const char*>(p.vals) + g * 128 + 96);
        return make_int2(kb[1 + (int)(idx - g * 3)], kb[0]);
    }
    return p.keys[idx];
}

// r / ng for a slot r inside a chunk (ng = 3 or 4).
ALLS_DEV int32_t group_of(const ListPool& p, int32_t r) {
    return p.ng == 4 ? (r >> 2) : (int32_t)(__umulhi((uint32_t)r, 0xAAAAAAABu) >> 1);
}

// 2D: records element e as the owner of the group at slot pos of row a (the
// group lies inside one chunk).  One 4-byte store per element row.
ALLS_DEV void list_store_group(const ListPool& p, int32_t a, int32_t pos, int32_t e) {
    const int64_t ch = list_chunk(p, a, pos);
    if (ch < 0) return;
    const int32_t r = pos - chunk_of(p, pos) * p.C;
    p.geid[ch * p.gpc + group_of(p, r)] = e;
}

#endif  // __HIPCC__

// ----------------------------------------------------------- host side ---
#ifdef __cplusplus
#include <string>
#include <vector>

namespace alls {

// Kernel launchers (defined in the .hip files).
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
// max_blocks > 0 caps the grid (persistent, grid-stride); 0 => one thread
// (2D) or one block (Hex20) per element.
void launch_insert_2d(const DevMesh& m, const ListPool& p, int mode, int block, int64_t max_blocks, hipStream_t s);
void launch_insert_hex20(const DevMesh& m, const ListPool& p, int f32, int64_t max_blocks, hipStream_t s);
// Pattern cache: build the scatter map (from the node -> element incidence);
// value-only and coloured scatter.
void launch_build_map(const DevMesh& m, int kind, const void* row_ptr, int i64, const int32_t* col_ind,
                      const int64_t* inc_off, const int32_t* inc_list, int32_t* map, int32_t* misc, hipStream_t s);
// atomic = 1: FP64 atomics (all elements at once); 0: plain "+=" (one colour).
void launch_scatter_2d(const DevMesh& m, int kind, const Scatter& sc, int atomic, int32_t* misc, hipStream_t s);
void launch_gather_2d(const DevMesh& m, int kind, const Scatter& sc, const int64_t* inc_off, const int32_t* inc_list,
                      int32_t* misc, hipStream_t s);
// Hex20 scatter: mode 1 = FP64 atomics, 0 = plain "+=" (one colour),
// 3 = write each block to stage[e*400 + I*20 + J] (value-only path).
void launch_scatter_hex20(const DevMesh& m, const Scatter& sc, int mode, int f32, int32_t* misc, hipStream_t s);
// Hex20 value-only path: the largest column count of a node row (atomicMax
// into *out), and the per-row ordered gather of element-order staged blocks.
void launch_row_umax(int64_t n_nodes, const void* rp, int i64, int32_t* out, hipStream_t s);
int32_t pair_gather_max_cols();
void launch_pair_gather(int64_t n_nodes, const int32_t* conn, const void* rp, int i64, const int32_t* rank_map,
                        const int64_t* inc_off, const int32_t* inc_list, const void* stage, int f32, int32_t umax,
                        double* val, hipStream_t s);
void launch_fill_f64(double* p, int64_t n, double v, hipStream_t s);
// Colouring: node -> element incidence, JPL rounds, conflict check, buckets.
void launch_incidence(const DevMesh& m, int32_t* cnt, int64_t* off, int32_t* cursor, int32_t* list, int64_t* tiles,
                      hipStream_t s);
void launch_jpl_round(const DevMesh& m, const int64_t* off, const int32_t* list, const int32_t* cin, int32_t* cout,
                      int32_t* misc, hipStream_t s);
void launch_conflicts(const DevMesh& m, const int64_t* off, const int32_t* list, const int32_t* col, int32_t* misc,
                      hipStream_t s);
// nc: number of colours (every col[e] lies in [0, nc)).
void launch_color_count(const int32_t* col, int64_t n, int nc, int32_t* cnt, hipStream_t s);
void launch_color_bucket(const int32_t* col, int64_t n, int nc, const int64_t* coff, int32_t* cursor,
                         int32_t* elist, hipStream_t s);
// One source mesh of a batch (device arrays) and its place in the packed mesh.
struct PackMesh {
    const double* coords;
    const int32_t* conn;
    const uint8_t* types;      // or null: every element is `type`
    const double* E_elem;      // or null: every element has E
    double E;
    int64_t n_nodes, n_elems, node_off, elem_off;
    int32_t npe;               // conn stride of the source
    int32_t type;
};
// Packs nm device meshes into coords/conn (and types, E when non-null).
void launch_pack(const PackMesh* pm, int nm, int dim, int npe, int64_t n_nodes, int64_t n_elems, double* coords,
                 int32_t* conn, uint8_t* types, double* E, hipStream_t s);
// One row/nonzero range of a block-diagonal batch result, copied into its
// own CSR with local row and column numbers.
struct SplitPart {
    int64_t row_off, nnz_off, n_rows, nnz;
    int32_t* row_ptr;           // int32 [n_rows+1]
    int32_t* col_ind;
    double* val;
};
void launch_split(const SplitPart* parts, int n, int64_t total_rows, int64_t total_nnz, const void* row_ptr,
                  int i64, const int32_t* col_ind, const double* val, hipStream_t s);
// out[i] = row_ptr[idx[i]] as int64.
void launch_gather_rp(const void* row_ptr, int i64, const int64_t* idx, int n, int64_t* out, hipStream_t s);

// Consolidation.  small_group = lanes per row of the register kernel (0 =>
// every row goes to the wave-per-row LDS kernel).
void launch_count(const ListPool& p, int small_group, int f32vals, int32_t* u, int32_t* big_list,
                  hipStream_t s);
void launch_scan_total(const int32_t* u, int64_t n, int64_t* tiles, int64_t ntiles, hipStream_t s);
void launch_scan_apply(const int32_t* u, int64_t n, const int64_t* tiles, int dpn, void* row_ptr,
                       int i64, hipStream_t s);
// max_big: most entries in a row of the wave-per-row kernel (MISC_MAX_BIG).
// big_list: the count pass's long-row list; when every row is long (3D
// lists) the write pass reuses it for the rows it defers.
// val: double[nnz], or float[nnz] when val_f32 (3D lists only; sums stay FP64).
void launch_write(const ListPool& p, int small_group, int f32vals, int dpn, int32_t max_big, int32_t* big_list,
                  const void* row_ptr, int i64, int32_t* col_ind, void* val, int val_f32, hipStream_t s);
int64_t scan_tile_count(int64_t n);
// Fused count + scan + write of 2D meshes whose rows all hold <= G entries
// (tstat: fused_tile_count int64 tile states, zeroed; misc[MISC_TICKET] zero).
int64_t fused_tile_count(int64_t n_nodes, int G);
void launch_rows_fused(const ListPool& p, int G, int64_t* tstat, void* row_ptr, int i64, int32_t* col_ind,
                       double* val, hipStream_t s);
// Most list entries one node row may hold (consolidation LDS limit).
int32_t long_row_limit();

}  // namespace alls
#endif
