// insert_hex20.hip -- Hex20 3D elasticity: MFMA element integration fused
// with A-LLS list insertion (gfx950).
//
// Replaces computeElementStiffnessHex20 (Hex:130-195) and
// assembleGlobalStiffnessKernelALLSFixedArray (Hex:208-257).
//
// For isotropic elasticity the 3x3 block of node pair (I,J) is
//   K_IJ = E * (lam * M_IJ + mu * M_IJ^T + mu * tr(M_IJ) * I3),
//   M[(I,p),(J,q)] = sum_g w_g * dN_I/dx_p(g) * dN_J/dx_q(g),
// the same sum the reference writes out term by term (Hex:173-191).  M is
// the 60x60 product X^T (W X) with X the 27x60 matrix of Gauss-point
// gradients, so one wavefront per element forms X in LDS (54 lanes: a Gauss
// point and half of the nodes each) and contracts it on the matrix cores:
// v_mfma_f64_16x16x4f64 (FP64 path) or v_mfma_f32_16x16x4f32 (FP32 path),
// over the 10 upper-triangle 16x16 tiles of the symmetric 64x64 (padded)
// product, 7 k-steps of 4 Gauss points.  M is kept in LDS as its 210 upper
// 3x3 node blocks, K is formed from them and every row node's 20 entries are
// appended to its list with one reservation atomic.  With chunks of whole
// 20-slot strips the lists carry the strip's element id once (its columns
// are the element's nodes) instead of a key per entry.
#include "hex20_block.h"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Gauss rule exactly as the reference writes it (Hex:45-46).
constexpr double HGP[3] = {-0.774596669241483, 0.0, 0.774596669241483};
constexpr double HGW[3] = {0.555555555555556, 0.888888888888889, 0.555555555555556};
// Natural coordinates of the 20 nodes (Hex:97-99).
constexpr double HXI[20] = {-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1, -1};
constexpr double HETA[20] = {-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1};
constexpr double HZETA[20] = {-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, 0, 0, 0, 0};

// LDS of one element (one wave), two phases over the same words:
//   gradients, transposed and scaled: Xt[P * XT + g] = sqrt|w_g| *
//   dN_n/dx_p at Gauss point g, at position P = 16 (n / 5) + 3 (n % 5) + p,
//   so that each 16-wide MFMA tile holds five whole nodes (positions 15, 31,
//   47 and 63 are never written: they only reach output rows and columns
//   that are discarded; the k-padding row g = 27 is masked at the load);
//   M as its 210 upper 3x3 node blocks, pair (I <= J) at 9 * pair_index(I, J),
//   row-major inside the block, then 64 words where lanes park the product
//   entries that belong to no stored block (branch-free stores).
constexpr int XT = 29;                  // Gauss points 27, padded to 28 k-rows, +1 against bank conflicts
constexpr int NPAIR = 210;              // node pairs I <= J
constexpr int MB_DUMP = NPAIR * 9;      // first of the 64 parking words
constexpr int HEX20_STAGE_ROWS_F32 = 3;   // node rows of list values staged per pass
constexpr int HEX20_STAGE_ROWS_F64 = 2;
constexpr int LDS_ELEMS = MB_DUMP + 64 > 64 * XT ? MB_DUMP + 64 : 64 * XT;

// Index of node pair (I, J), I <= J, in the packed upper triangle.

// dN/d(xi,eta,zeta) of node n at a Gauss point (Hex:95-128).
constexpr void hex20_dshape(int n, double xi, double eta, double zeta, double& dx, double& de, double& dz) {
    const double a = HXI[n], b = HETA[n], c = HZETA[n];
    if (n < 8) {
        dx = 0.125 * a * (1.0 + b * eta) * (1.0 + c * zeta) * (2.0 * a * xi + b * eta + c * zeta - 1.0);
        de = 0.125 * b * (1.0 + a * xi) * (1.0 + c * zeta) * (a * xi + 2.0 * b * eta + c * zeta - 1.0);
        dz = 0.125 * c * (1.0 + a * xi) * (1.0 + b * eta) * (a * xi + b * eta + 2.0 * c * zeta - 1.0);
    } else if (n >= 16) {
        dx = 0.25 * a * (1.0 + b * eta) * (1.0 - zeta * zeta);
        de = 0.25 * b * (1.0 + a * xi) * (1.0 - zeta * zeta);
        dz = -0.5 * zeta * (1.0 + a * xi) * (1.0 + b * eta);
    } else if ((n & 1) == 0) {  // 8, 10, 12, 14: xi_n = 0
        dx = -0.5 * xi * (1.0 + b * eta) * (1.0 + c * zeta);
        de = 0.25 * b * (1.0 - xi * xi) * (1.0 + c * zeta);
        dz = 0.25 * c * (1.0 - xi * xi) * (1.0 + b * eta);
    } else {                    // 9, 11, 13, 15: eta_n = 0
        dx = 0.25 * a * (1.0 - eta * eta) * (1.0 + c * zeta);
        de = -0.5 * eta * (1.0 + a * xi) * (1.0 + c * zeta);
        dz = 0.25 * c * (1.0 + a * xi) * (1.0 - eta * eta);
    }
}

// Natural derivatives dN/d(xi, eta, zeta) of the 20 nodes at the 27 Gauss
// points (g = 9 i + 3 j + k for xi_i, eta_j, zeta_k), evaluated at compile
// time and laid out per gradient lane: lane (g, h), which takes nodes
// 10h .. 10h+9, reads its 30 values from [(27 h + g) * 30].  They do not
// depend on the element; 13 KB, cache-resident.
struct alignas(16) DnTable {
    double v[2 * 27 * 30];
};
constexpr DnTable make_dn_table() {
    DnTable t{};
    for (int h = 0; h < 2; ++h)
        for (int g = 0; g < 27; ++g)
            for (int k = 0; k < 10; ++k) {
                double dx = 0, de = 0, dz = 0;
                hex20_dshape(10 * h + k, HGP[g / 9], HGP[(g / 3) % 3], HGP[g % 3], dx, de, dz);
                const int o = (27 * h + g) * 30 + 3 * k;
                t.v[o] = dx;
                t.v[o + 1] = de;
                t.v[o + 2] = dz;
            }
    return t;
}
__constant__ DnTable HDN = make_dn_table();

// Equal row nodes among the element's 20 node lanes: rank = lanes below
// with the same node, cnt = lanes with it, lead = the lowest of them.
// Lanes 20..63 keep (0, 1, lane).
__device__ __forceinline__ void hex20_match(int32_t n, int lane, int32_t* rk, int32_t* cnt, int32_t* lead) {
    int32_t r = 0, c = 0, l = lane;
#pragma unroll
    for (int j = 0; j < 20; ++j) {
        const int32_t nj = __builtin_amdgcn_readlane(n, j);
        if (nj == n) {
            ++c;
            r += j < lane ? 1 : 0;
            l = j < l ? j : l;
        }
    }
    if (lane < 20) {
        *rk = r;
        *cnt = c;
        *lead = l;
    }
}

__device__ __forceinline__ f64x4 mfma(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Acc;
template <> struct Acc<double> { typedef f64x4 type; };
template <> struct Acc<float> { typedef f32x4 type; };
// Row of the 16x16 output held in register q of lane (li + 16*lk); the
// column is li.  Measured on gfx950: the FP64 16x16x4 instruction
// interleaves rows across lane groups (row 4q + lk), the FP32 one keeps four
// consecutive rows per lane group (row 4lk + q).  A and B take A[i][k] and
// B[k][j] from lane i + 16k / j + 16k in both.
template <typename T> struct OutRow;
template <> struct OutRow<double> {
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * q + lk; }
};
template <> struct OutRow<float> {
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

// Rows of list values staged per pass of the insertion (LDS: rows * 720 B
// in FP32, rows * 1440 B in FP64).
template <typename T> struct Hex20Stage;
template <> struct Hex20Stage<float> { static constexpr int rows = HEX20_STAGE_ROWS_F32; };
template <> struct Hex20Stage<double> { static constexpr int rows = HEX20_STAGE_ROWS_F64; };

// One 16-byte piece of a staged row strip from LDS to the list pool,
// non-temporal: the merge reads it back once, and the pieces of a wave's
// store instruction are contiguous (80^3, one box: FP64 insertion 7.3 ->
// 5.0-5.7 ms and FP32 4.4 -> 2.9-3.0 ms against plain stores, two rows per
// pass in FP64 and three in FP32; 5 and 10 rows measured slower).
template <typename T>
struct Piece {
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};
template <typename T>
__device__ __forceinline__ typename Piece<T>::type get_piece(const T* src) {
    return *reinterpret_cast<const typename Piece<T>::type*>(src);
}
template <typename T>
__device__ __forceinline__ void put_piece(T* dst, typename Piece<T>::type v) {
    __builtin_nontemporal_store(v, reinterpret_cast<typename Piece<T>::type*>(dst));
}

// EMIT 0: append to the A-LLS lists; 1: FP64-atomic scatter into a cached
// pattern; 2: plain "+=" scatter (one colour); 3: write each block to its
// precomputed staging slot (value-only re-assembly, no atomics).
// FP32 asks for 4 waves per SIMD (128 VGPRs, no spills; LDS allows 14 per
// CU), FP64 for 2 (150 VGPRs; 19 KB of LDS allows 8 per CU): FP32 insertion
// 2.95 -> 2.77 ms at 80^3 together with its early reservations (three
// interleaved rounds on one box).
template <typename T, int EMIT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) void k_hex20(
    DevMesh m, ListPool p, Scatter sc, int32_t* misc) {
    __shared__ __align__(16) T lds[LDS_ELEMS];
    __shared__ double xyz[60];
    __shared__ T wg[28];
    __shared__ int32_t nodes[20];
    __shared__ int32_t posv[20];
    __shared__ int64_t rbase[20];   // list index of each row group's first slot (-1: spans chunks, -2: none)
    __shared__ int32_t bad;
    __shared__ __align__(16) T kst[EMIT == 0 || EMIT == 3 ? Hex20Stage<T>::rows * 180 : 4];

    const int lane = threadIdx.x;
    const int64_t count = EMIT == 0 ? m.n_elems : sc.count;
    // Grid-stride over the elements, each XCD over its own slab.
    const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, count);
    for (int64_t idx = xr.first; idx < xr.end; idx += xr.step) {
        // The lane id, opaque to the compiler inside the loop: everything the
        // loop derives from it (table loads, LDS indices) would otherwise be
        // hoisted out and held in registers across the whole loop, which
        // spills.
        int ln = lane;
        __asm__ volatile("" : "+v"(ln));
        // Gradient phase: lane (g, h) takes Gauss point g < 27 and nodes
        // 10h .. 10h+9 (lanes with g >= 27 compute point 26 and store nothing).
        const int g = ln & 31, h = ln >> 5;
        const int gq = g < 27 ? g : 26;
        const double2* dn_src = reinterpret_cast<const double2*>(HDN.v + (27 * h + gq) * 30);
        const int64_t e = (EMIT != 0 && sc.elist) ? (int64_t)sc.elist[idx] : idx;
        if (lane == 0) bad = 0;
        __syncthreads();
        // The reservation atomics go out with the coordinate loads, so their
        // round trip overlaps the integration (an element with a bad node
        // makes the whole call fail, so its reserved slots are never read).
        constexpr bool early = EMIT == 0;
        const int32_t n = lane < 20 ? m.conn[e * m.npe + lane] : -1;
        // This lane's natural derivatives (cache hits), in flight with the coordinates.
        double dn[10][3];
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            const double2 v = dn_src[k];
            dn[(2 * k) / 3][(2 * k) % 3] = v.x;
            dn[(2 * k + 1) / 3][(2 * k + 1) % 3] = v.y;
        }
        const bool valid = lane < 20 && n >= 0 && n < m.n_nodes;
        // Lanes naming the same row node (an element that lists a node
        // twice) share one reservation: the lowest such lane reserves 20
        // slots for each of them and lane order assigns the strips, so a
        // row's strips of one element lie in the element's row-node order.
        int32_t rk = 0, cnt = 1, lead = lane;
        if (EMIT == 0) hex20_match(n, lane, &rk, &cnt, &lead);
        int32_t pos = 0;
        if (lane < 20) {
            if (!valid) {
                bad = 1;
                nodes[lane] = 0;
            } else {
                nodes[lane] = n;
                const double cx = m.coords[3 * (int64_t)n + 0];
                const double cy = m.coords[3 * (int64_t)n + 1];
                const double cz = m.coords[3 * (int64_t)n + 2];
                if (early && rk == 0) pos = list_reserve_issue(p, n, 20 * cnt);
                xyz[lane * 3 + 0] = cx;
                xyz[lane * 3 + 1] = cy;
                xyz[lane * 3 + 2] = cz;
            }
        }
        __syncthreads();
        if (bad || (m.types && m.types[e] != 2)) {
            if (lane == 0) st_agent(&misc[MISC_ERR + ERR_INVALID_MESH], 1);
            // Reservations already taken still publish the chunks they own:
            // later elements of the same rows wait for those chunks.
            if (early && valid && rk == 0) list_own_chunks(p, n, pos, 20 * cnt);
            __syncthreads();
            continue;
        }

        // Jacobian at Gauss point g: each half sums its 10 nodes, then the
        // halves add (in either lane the same two terms, so the same bits).
        T* Xt = lds;
        {
            double Jp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 10; ++t) {
                const int nd = 10 * h + t;
                const double px = xyz[3 * nd], py = xyz[3 * nd + 1], pz = xyz[3 * nd + 2];
                Jp[0] += dn[t][0] * px; Jp[1] += dn[t][0] * py; Jp[2] += dn[t][0] * pz;
                Jp[3] += dn[t][1] * px; Jp[4] += dn[t][1] * py; Jp[5] += dn[t][1] * pz;
                Jp[6] += dn[t][2] * px; Jp[7] += dn[t][2] * py; Jp[8] += dn[t][2] * pz;
            }
            double J[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) J[i] = Jp[i] + __shfl_xor(Jp[i], 32, 64);
            double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                         J[2] * (J[3] * J[7] - J[4] * J[6]);
            if (fabs(det) < 1e-15) det = 1e-15;  // Hex:156
            const double r = 1.0 / det;
            const double i0 = (J[4] * J[8] - J[5] * J[7]) * r, i1 = (J[2] * J[7] - J[1] * J[8]) * r;
            const double i2 = (J[1] * J[5] - J[2] * J[4]) * r, i3 = (J[5] * J[6] - J[3] * J[8]) * r;
            const double i4 = (J[0] * J[8] - J[2] * J[6]) * r, i5 = (J[2] * J[3] - J[0] * J[5]) * r;
            const double i6 = (J[3] * J[7] - J[4] * J[6]) * r, i7 = (J[1] * J[6] - J[0] * J[7]) * r;
            const double i8 = (J[0] * J[4] - J[1] * J[3]) * r;
            if (g < 27) {
                // Stored scaled by sqrt|w|, w = Gauss weight * det, with the
                // sign of w kept apart: M = Z^T (sign(w) Z) takes one rounding
                // per operand value, as a separate W X would, in one array.
                const double w = HGW[g / 9] * HGW[(g / 3) % 3] * HGW[g % 3] * det;
                const double sw = sqrt(fabs(w));
#pragma unroll
                for (int t = 0; t < 10; ++t) {
                    const int c = 32 * h + 16 * (t / 5) + 3 * (t % 5);   // node 10h + t
                    const double gx = i0 * dn[t][0] + i1 * dn[t][1] + i2 * dn[t][2];
                    const double gy = i3 * dn[t][0] + i4 * dn[t][1] + i5 * dn[t][2];
                    const double gz = i6 * dn[t][0] + i7 * dn[t][1] + i8 * dn[t][2];
                    Xt[(c + 0) * XT + g] = (T)(gx * sw);
                    Xt[(c + 1) * XT + g] = (T)(gy * sw);
                    Xt[(c + 2) * XT + g] = (T)(gz * sw);
                }
                if (h == 0) wg[g] = w < 0.0 ? (T)-1 : (T)1;
            }
        }
        __syncthreads();

        // M = X^T (W X) = Z^T (S Z) on the matrix cores, Z = sqrt|W| X and S
        // = sign(W): tiles (tr <= tc) of 16x16, A = Z^T and B = S Z from the
        // same LDS word; k-row 27 (padding) is zero in both.
        typedef typename Acc<T>::type acc_t;
        acc_t acc[10];
    #pragma unroll
        for (int t = 0; t < 10; ++t) acc[t] = acc_t{0, 0, 0, 0};
        const int li = ln & 15, lk = ln >> 4;
    #pragma unroll
        for (int s = 0; s < 7; ++s) {
            const bool pad = s == 6 && lk == 3;
            const T wk = pad ? (T)0 : wg[4 * s + lk];
            T av[4], bv[4];
    #pragma unroll
            for (int t = 0; t < 4; ++t) {
                const T x = Xt[(16 * t + li) * XT + 4 * s + lk];
                av[t] = pad ? (T)0 : x;
                bv[t] = av[t] * wk;
            }
            int t = 0;
    #pragma unroll
            for (int tr = 0; tr < 4; ++tr)
    #pragma unroll
                for (int tc = tr; tc < 4; ++tc) {
                    acc[t] = mfma(av[tr], bv[tc], acc[t]);
                    ++t;
                }
        }
        __syncthreads();   // Xt is overwritten by M below

        // The tiles into the packed upper blocks of M.  Entry (r16, c16) of
        // tile (tr, tc) is entry (rr, cc) = (r16 % 3, c16 % 3) of node block
        // (I, J) = (5 tr + r16 / 3, 5 tc + c16 / 3) (r16, c16 < 15), kept when
        // I <= J, at 9 pair_index(I, J) + 3 rr + cc; no block crosses a tile.
        T* Mb = lds;
        {
            const int jl = (li * 11) >> 5, cc = li - 3 * jl;   // li / 3, li % 3 for li < 16
            int t = 0;
    #pragma unroll
            for (int tr = 0; tr < 4; ++tr)
    #pragma unroll
                for (int tc = tr; tc < 4; ++tc) {
    #pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r16 = OutRow<T>::row(lk, q);
                        const int il = (r16 * 11) >> 5, rr = r16 - 3 * il;
                        const int I = 5 * tr + il, J = 5 * tc + jl;
                        const bool keep = li < 15 && r16 < 15 && (tr < tc || il <= jl);
                        const int at = 9 * pair_index(I, J) + 3 * rr + cc;
                        Mb[keep ? at : MB_DUMP + lane] = acc[t][q];
                    }
                    ++t;
                }
        }
        const double E = m.E_elem ? m.E_elem[e] : m.E;
        const double nu = m.nu;
        const T lam = (T)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
        const T mu = (T)(E / (2.0 * (1.0 + nu)));
        if (EMIT == 0) {
            // The 20 slots reserved in each of the element's 20 node rows.
            if (lane < 20 && rk == 0) {
                list_note_long(p, pos, 20 * cnt);
                list_own_chunks(p, nodes[lane], pos, 20 * cnt);
            }
            pos = __shfl(pos, lead, 64) + 20 * rk;   // this lane's strip of the shared reservation
            if (lane < 20) {
                posv[lane] = pos;
                // The group's chunk, found once per row instead of once per
                // entry (an overflow chunk's directory word is an L2 access).
                const int32_t k0 = chunk_of(p, pos);
                int64_t b = -1;
                if (k0 == chunk_of(p, pos + 19)) {
                    const int64_t ch = list_chunk(p, nodes[lane], pos);
                    b = ch < 0 ? -2 : ch * p.C + (pos - k0 * p.C);
                }
                rbase[lane] = b;
                // Segment lists: the strip's element id, one store per row
                // (a strip never crosses a chunk there, so b >= 0 unless
                // the chunk could not be had, which fails the call).
                if (p.ng == SEG20 && b >= 0) p.geid[b / SEG20] = (int32_t)e;
            }
        } else if (EMIT == 3) {
            // Element-order staging: row I's 20 blocks at entries e*400 + 20 I.
            if (lane < 20) rbase[lane] = e * 400 + 20 * lane;
        }
        __syncthreads();

        if constexpr (EMIT == 0 || EMIT == 3) {
            // Rows of list values in passes of RS rows: the K blocks of the
            // pass go to LDS in list order (a row's 20 entries are consecutive
            // list slots), then consecutive lanes copy consecutive 16-byte
            // pieces of each row strip, so every store instruction covers
            // whole runs of memory.  A row whose strip is not 16-byte aligned
            // in the pool (chunks that are not a multiple of 4 or 2 slots, or
            // a strip across chunks) is written block by block instead.  The
            // value-only path (EMIT 3) writes the same strips, in element
            // order, to the staging array (always aligned, no keys).
            constexpr int RS = Hex20Stage<T>::rows, VPP = 16 / (int)sizeof(T), PPR = 180 / VPP;
            constexpr int64_t ALIGN = sizeof(T) == 4 ? 4 : 2;   // slots per 16 bytes of values
            // Fixed trip counts per pass (KIT K-formation and CIT copy-out
            // steps of 64 lanes), so that the compiler issues a pass's LDS
            // reads together instead of one round trip per 16-byte piece.
            constexpr int KIT = (RS * 20 + 63) / 64, CIT = (RS * PPR + 63) / 64;
            T* dst = static_cast<T*>(EMIT == 0 ? p.vals : sc.stage);
            for (int r0 = 0; r0 < 20; r0 += RS) {
                const int nr = 20 - r0 < RS ? 20 - r0 : RS;
    #pragma unroll
                for (int it = 0; it < KIT; ++it) {
                    const int q = ln + 64 * it;
                    if (q >= nr * 20) break;
                    const int ri = q / 20, I = r0 + ri, Jn = q - 20 * ri;
                    T kb[3][3];
                    k_block(Mb, I, Jn, lam, mu, kb);
                    int64_t id = rbase[I];
                    if (id == -2) continue;
                    if (EMIT == 3 || (id >= 0 && id % ALIGN == 0)) {
                        T* st = kst + ri * 180 + Jn * 9;
    #pragma unroll
                        for (int v = 0; v < 9; ++v) st[v] = kb[v / 3][v % 3];
                        if (EMIT == 0 && p.ng != SEG20) p.keys[id + Jn] = make_int2(nodes[Jn], (int32_t)e);
                        continue;
                    }
                    if (id >= 0) {
                        id += Jn;
                    } else {   // the strip spans two chunks (C not a multiple of 20)
                        const int32_t s = posv[I] + Jn;
                        const int64_t ch = list_chunk(p, nodes[I], s);
                        if (ch < 0) continue;
                        id = ch * p.C + (s - chunk_of(p, s) * p.C);
                    }
                    T* d = static_cast<T*>(p.vals) + id * 9;
    #pragma unroll
                    for (int r = 0; r < 3; ++r)
    #pragma unroll
                        for (int c = 0; c < 3; ++c) d[r * 3 + c] = kb[r][c];
                    if (p.ng != SEG20) p.keys[id] = make_int2(nodes[Jn], (int32_t)e);
                }
                __syncthreads();
                // The pass's row bases, read once (uniform per row), and all
                // of its pieces read from LDS (index clamped into the stage)
                // before any is stored.
                int64_t rbp[RS];
    #pragma unroll
                for (int r = 0; r < RS; ++r) rbp[r] = rbase[r0 + r < 20 ? r0 + r : 19];
                typename Piece<T>::type pv[CIT];
    #pragma unroll
                for (int it = 0; it < CIT; ++it) {
                    const int pc = min(ln + 64 * it, RS * PPR - 1);
                    pv[it] = get_piece(kst + pc * VPP);   // piece pc of the stage (rows are 180 values, PPR pieces)
                }
    #pragma unroll
                for (int it = 0; it < CIT; ++it) {
                    const int pc = ln + 64 * it;
                    const int ri = pc / PPR, off = pc - ri * PPR;
                    int64_t rb = rbp[0];
    #pragma unroll
                    for (int r = 1; r < RS; ++r) rb = ri == r ? rbp[r] : rb;
                    if (pc < nr * PPR && rb >= 0 && rb % ALIGN == 0) put_piece(dst + rb * 9 + off * VPP, pv[it]);
                }
                __syncthreads();
            }
        } else {
        for (int pr = ln; pr < 400; pr += 64) {
            const int I = pr / 20, Jn = pr % 20;
            T kb[3][3];
            k_block(Mb, I, Jn, lam, mu, kb);
            {
                const int64_t a = nodes[I];
                const int64_t rs = rp_at(sc.row_ptr, sc.i64, 3 * a);
                const int64_t rowlen = rp_at(sc.row_ptr, sc.i64, 3 * a + 1) - rs;
                double* v = sc.val + rs + 3 * (int64_t)sc.map[e * 400 + pr];
    #pragma unroll
                for (int r = 0; r < 3; ++r)
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double* q = v + r * rowlen + c;
                        if (EMIT == 1) atomicAdd(q, (double)kb[r][c]);
                        else *q += (double)kb[r][c];
                    }
            }
        }
        }
        __syncthreads();   // LDS is refilled by the next element
    }
}

}  // namespace

namespace alls {

// Grid of k_hex20: one block per element, or max_blocks > 0 (persistent).
static int64_t hex20_grid(int64_t count, int64_t max_blocks) {
    return max_blocks > 0 && count > max_blocks ? max_blocks : count;
}

void launch_insert_hex20(const DevMesh& m, const ListPool& p, int f32, int64_t max_blocks, hipStream_t s) {
    if (m.n_elems <= 0) return;
    const Scatter none{};
    const int64_t g = hex20_grid(m.n_elems, max_blocks);
    if (f32)
        hipLaunchKernelGGL((k_hex20<float, 0>), dim3((unsigned)g), dim3(64), 0, s, m, p, none, p.misc);
    else
        hipLaunchKernelGGL((k_hex20<double, 0>), dim3((unsigned)g), dim3(64), 0, s, m, p, none, p.misc);
}

void launch_scatter_hex20(const DevMesh& m, const Scatter& sc, int mode, int f32, int32_t* misc, hipStream_t s) {
    if (sc.count <= 0) return;
    const ListPool none{};
    const dim3 g((unsigned)hex20_grid(sc.count, 0));
#define ALLS_HX(T, E) hipLaunchKernelGGL((k_hex20<T, E>), g, dim3(64), 0, s, m, none, sc, misc)
    if (f32) {
        if (mode == 1) ALLS_HX(float, 1);
        else if (mode == 3) ALLS_HX(float, 3);
        else ALLS_HX(float, 2);
    } else {
        if (mode == 1) ALLS_HX(double, 1);
        else if (mode == 3) ALLS_HX(double, 3);
        else ALLS_HX(double, 2);
    }
#undef ALLS_HX
}

}  // namespace alls
