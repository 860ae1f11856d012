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
// gradients, so one wavefront per element forms X in LDS and contracts it
// on the matrix cores: v_mfma_f64_16x16x4f64 (FP64 path) or
// v_mfma_f32_16x16x4f32 (FP32 path), over the 10 upper-triangle 16x16 tiles
// of the symmetric 64x64 (padded) product, 7 k-steps of 4 Gauss points.
// K is then formed from M in LDS and every row node's 20 entries are
// appended to its list with one reservation atomic.
#include "alls_internal.h"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// A block value into the lists or the element-order stage, read once by a
// later kernel: non-temporal in FP32 (insertion 4.59 -> 4.35 ms, value-only
// 7.35 -> 7.08 ms at 80^3), plain in FP64, where the 8-byte non-temporal
// stores did not combine and the insertion took 22.5 ms instead of 7.9.
template <typename T>
__device__ __forceinline__ void put_block_value(T* dst, T x) {
    if constexpr (sizeof(T) == 4) __builtin_nontemporal_store(x, dst);
    else *dst = x;
}

// FP64 blocks of lanes 2m and 2m+1, which hold consecutive entries of the
// element-order stage (9 doubles each; the pair starts 16-byte aligned): the pair's
// 144 bytes go out as nine 16-byte stores, five from the even lane (its
// block and the odd lane's first value) and four from the odd lane, instead
// of eighteen 8-byte stores.  `dst` is the lane's own entry; both lanes of a
// pair must be active.
__device__ __forceinline__ void put_block_pair(double* dst, const double (&kb)[3][3], bool even) {
    const double first_of_odd = __shfl_xor(kb[0][0], 1, 64);
    double2* d = reinterpret_cast<double2*>(even ? dst : dst + 1);
    if (even) {
        d[0] = make_double2(kb[0][0], kb[0][1]);
        d[1] = make_double2(kb[0][2], kb[1][0]);
        d[2] = make_double2(kb[1][1], kb[1][2]);
        d[3] = make_double2(kb[2][0], kb[2][1]);
        d[4] = make_double2(kb[2][2], first_of_odd);
    } else {
        d[0] = make_double2(kb[0][1], kb[0][2]);
        d[1] = make_double2(kb[1][0], kb[1][1]);
        d[2] = make_double2(kb[1][2], kb[2][0]);
        d[3] = make_double2(kb[2][1], kb[2][2]);
    }
}
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Gauss rule exactly as the reference writes it (Hex:45-46).
__constant__ double HGP[3] = {-0.774596669241483, 0.0, 0.774596669241483};
__constant__ double HGW[3] = {0.555555555555556, 0.888888888888889, 0.555555555555556};
// Natural coordinates of the 20 nodes (Hex:97-99).
__constant__ double HXI[20] = {-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1, -1};
__constant__ double HETA[20] = {-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1};
__constant__ double HZETA[20] = {-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, 0, 0, 0, 0};

constexpr int KS = 28;   // Gauss rows padded 27 -> 28 (7 k-steps of 4)
constexpr int XS = 64;   // gradient columns padded 60 -> 64
constexpr int MS = 61;   // row stride of M in LDS

// dN/d(xi,eta,zeta) of node n at a Gauss point (Hex:95-128).
__device__ __forceinline__ void hex20_dshape(int n, double xi, double eta, double zeta, double& dx, double& de,
                                             double& dz) {
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

// EMIT 0: append to the A-LLS lists; 1: FP64-atomic scatter into a cached
// pattern; 2: plain "+=" scatter (one colour); 3: write each block to its
// precomputed staging slot (value-only re-assembly, no atomics).
template <typename T, int EMIT>
__global__ __launch_bounds__(64, 2) void k_hex20(DevMesh m, ListPool p, Scatter sc, int32_t* misc) {
    constexpr int LDS_T = (2 * KS * XS > 60 * MS) ? 2 * KS * XS : 60 * MS;
    __shared__ T lds[LDS_T];
    __shared__ double xyz[60];
    __shared__ int32_t nodes[20];
    __shared__ int32_t posv[20];
    __shared__ int64_t rbase[20];   // list index of each row group's first slot (-1: spans chunks, -2: none)
    __shared__ int32_t bad;

    const int lane = threadIdx.x;
    const int64_t count = EMIT == 0 ? m.n_elems : sc.count;
    // Grid-stride over the elements (a grid sized to the device, or one block per element).
    for (int64_t idx = xcd_block(blockIdx.x, gridDim.x); idx < count; idx += gridDim.x) {
        const int64_t e = (EMIT != 0 && sc.elist) ? (int64_t)sc.elist[idx] : idx;
        if (lane == 0) bad = 0;
        __syncthreads();
        // FP64: the reservation atomics go out with the coordinate loads, so
        // their round trip overlaps the integration (80^3: 8.1 -> 7.8 ms; an
        // element with a bad node makes the whole call fail, so its reserved
        // slots are never read).  FP32 reserves after the contraction: the
        // register held across it cost 3.2 -> 3.6 ms there.
        constexpr bool early = EMIT == 0 && sizeof(T) == 8;
        const int32_t n = lane < 20 ? m.conn[e * m.npe + lane] : -1;
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

        // Gauss-point gradients: lane g < 27 fills row g of X and W*X.
        T* X = lds;
        T* Y = lds + KS * XS;
        if (lane < KS) {
            if (lane < 27) {
                const int gi = lane / 9, gj = (lane / 3) % 3, gk = lane % 3;
                const double xi = HGP[gi], eta = HGP[gj], zeta = HGP[gk];
                const double weight = HGW[gi] * HGW[gj] * HGW[gk];
                double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int n = 0; n < 20; ++n) {
                    double dx, de, dz;
                    hex20_dshape(n, xi, eta, zeta, dx, de, dz);
                    const double px = xyz[3 * n], py = xyz[3 * n + 1], pz = xyz[3 * n + 2];
                    J[0] += dx * px; J[1] += dx * py; J[2] += dx * pz;
                    J[3] += de * px; J[4] += de * py; J[5] += de * pz;
                    J[6] += dz * px; J[7] += dz * py; J[8] += dz * pz;
                }
                double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                             J[2] * (J[3] * J[7] - J[4] * J[6]);
                if (fabs(det) < 1e-15) det = 1e-15;  // Hex:156
                const double r = 1.0 / det;
                const double i0 = (J[4] * J[8] - J[5] * J[7]) * r, i1 = (J[2] * J[7] - J[1] * J[8]) * r;
                const double i2 = (J[1] * J[5] - J[2] * J[4]) * r, i3 = (J[5] * J[6] - J[3] * J[8]) * r;
                const double i4 = (J[0] * J[8] - J[2] * J[6]) * r, i5 = (J[2] * J[3] - J[0] * J[5]) * r;
                const double i6 = (J[3] * J[7] - J[4] * J[6]) * r, i7 = (J[1] * J[6] - J[0] * J[7]) * r;
                const double i8 = (J[0] * J[4] - J[1] * J[3]) * r;
                const double w = weight * det;
                T* xr = X + lane * XS;
                T* yr = Y + lane * XS;
                for (int n = 0; n < 20; ++n) {
                    double dx, de, dz;
                    hex20_dshape(n, xi, eta, zeta, dx, de, dz);
                    const double gx = i0 * dx + i1 * de + i2 * dz;
                    const double gy = i3 * dx + i4 * de + i5 * dz;
                    const double gz = i6 * dx + i7 * de + i8 * dz;
                    xr[3 * n + 0] = (T)gx; xr[3 * n + 1] = (T)gy; xr[3 * n + 2] = (T)gz;
                    yr[3 * n + 0] = (T)(w * gx); yr[3 * n + 1] = (T)(w * gy); yr[3 * n + 2] = (T)(w * gz);
                }
                for (int c = 60; c < XS; ++c) xr[c] = yr[c] = (T)0;
            } else {
                for (int c = 0; c < XS; ++c) X[lane * XS + c] = Y[lane * XS + c] = (T)0;
            }
        }
        __syncthreads();

        // M = X^T Y on the matrix cores: tiles (tr <= tc) of 16x16, A = X^T, B = Y.
        typedef typename Acc<T>::type acc_t;
        acc_t acc[10];
    #pragma unroll
        for (int t = 0; t < 10; ++t) acc[t] = acc_t{0, 0, 0, 0};
        const int li = lane & 15, lk = lane >> 4;
    #pragma unroll
        for (int s = 0; s < KS / 4; ++s) {
            T av[4], bv[4];
    #pragma unroll
            for (int t = 0; t < 4; ++t) {
                av[t] = X[(4 * s + lk) * XS + 16 * t + li];
                bv[t] = Y[(4 * s + lk) * XS + 16 * t + li];
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
        __syncthreads();   // X/Y are overwritten by M below

        // Scatter the tiles into M (row stride MS), mirroring the upper triangle.
        T* M = lds;
        {
            int t = 0;
    #pragma unroll
            for (int tr = 0; tr < 4; ++tr)
    #pragma unroll
                for (int tc = tr; tc < 4; ++tc) {
    #pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r = 16 * tr + OutRow<T>::row(lk, q), c = 16 * tc + li;
                        if (r < 60 && c < 60) {
                            M[r * MS + c] = acc[t][q];
                            if (tr != tc) M[c * MS + r] = acc[t][q];
                        }
                    }
                    ++t;
                }
        }
        const double E = m.E_elem ? m.E_elem[e] : m.E;
        const double nu = m.nu;
        const T lam = (T)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
        const T mu = (T)(E / (2.0 * (1.0 + nu)));
        if (EMIT == 0) {
            // Reserve 20 slots in each of the element's 20 node rows.
            if (lane < 20 && rk == 0) {
                if (!early) pos = list_reserve_issue(p, nodes[lane], 20 * cnt);
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
            }
        }
        __syncthreads();

        for (int pr = lane; pr < 400; pr += 64) {
            const int I = pr / 20, Jn = pr % 20;
            T mb[3][3];
    #pragma unroll
            for (int r = 0; r < 3; ++r)
    #pragma unroll
                for (int c = 0; c < 3; ++c) mb[r][c] = M[(3 * I + r) * MS + 3 * Jn + c];
            const T tr = mb[0][0] + mb[1][1] + mb[2][2];
            T kb[3][3];
    #pragma unroll
            for (int r = 0; r < 3; ++r)
    #pragma unroll
                for (int c = 0; c < 3; ++c) kb[r][c] = lam * mb[r][c] + mu * mb[c][r] + (r == c ? mu * tr : (T)0);
            if (EMIT == 0) {
                int64_t id = rbase[I];
                if (id == -2) continue;
                if (id >= 0) {
                    id += Jn;
                } else {   // the group spans two chunks (C not a multiple of 20)
                    const int32_t s = posv[I] + Jn;
                    const int64_t ch = list_chunk(p, nodes[I], s);
                    if (ch < 0) continue;
                    id = ch * p.C + (s - chunk_of(p, s) * p.C);
                }
                T* dst = static_cast<T*>(p.vals) + id * 9;
                // FP64 rows whose group starts at an even slot (every row with
                // 160-slot chunks): lane pairs (2m, 2m+1) hold consecutive
                // blocks of one row (20 is even), 144 bytes from a 16-byte
                // boundary, written as nine 16-byte stores (80^3: 8.1 -> 7.55
                // ms; at 80-slot chunks, where groups sat across chunks, the
                // pairs measured 2.5% slower).  Both lanes of a pair take the
                // same branch: the test depends on the row only.
                bool paired = false;
                if constexpr (sizeof(T) == 8) {
                    if (rbase[I] >= 0 && (rbase[I] & 1) == 0) {
                        put_block_pair(dst, kb, (pr & 1) == 0);
                        paired = true;
                    }
                }
                if (!paired) {
    #pragma unroll
                    for (int r = 0; r < 3; ++r)
    #pragma unroll
                        for (int c = 0; c < 3; ++c) put_block_value(dst + r * 3 + c, kb[r][c]);
                }
                p.keys[id] = make_int2(nodes[Jn], (int32_t)e);
            } else if (EMIT == 3) {
                T* dst = static_cast<T*>(sc.stage) + (e * 400 + pr) * 9;
                if constexpr (sizeof(T) == 8) {
                    put_block_pair(dst, kb, (pr & 1) == 0);   // pr and lane have the same parity
                } else {
    #pragma unroll
                    for (int r = 0; r < 3; ++r)
    #pragma unroll
                        for (int c = 0; c < 3; ++c) put_block_value(dst + r * 3 + c, kb[r][c]);
                }
            } else {
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
        __syncthreads();   // LDS is refilled by the next element
    }
}

}  // namespace

namespace alls {

void launch_insert_hex20(const DevMesh& m, const ListPool& p, int f32, int64_t max_blocks, hipStream_t s) {
    if (m.n_elems <= 0) return;
    const Scatter none{};
    int64_t g = m.n_elems;
    if (max_blocks > 0 && g > max_blocks) g = max_blocks;
    if (f32)
        hipLaunchKernelGGL((k_hex20<float, 0>), dim3((unsigned)g), dim3(64), 0, s, m, p, none, p.misc);
    else
        hipLaunchKernelGGL((k_hex20<double, 0>), dim3((unsigned)g), dim3(64), 0, s, m, p, none, p.misc);
}

void launch_scatter_hex20(const DevMesh& m, const Scatter& sc, int mode, int f32, int32_t* misc, hipStream_t s) {
    if (sc.count <= 0) return;
    const ListPool none{};
    const dim3 g((unsigned)sc.count);
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
