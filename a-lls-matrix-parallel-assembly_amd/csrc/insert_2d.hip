// insert_2d.hip -- Q4 / T3 / mixed plane-stress element integration fused
// with A-LLS list insertion (gfx950).
//
// Replaces assembleGlobalStiffnessKernelALLS (Acc:394-468, T3:227-290,
// Mix:248-312).  The reference walks each row's linked list per scalar
// contribution and merges equal columns with FP64 atomics.  Here one thread
// integrates one element, keeps the Gauss-point moments
//   Sxx_ab = sum w bx_a bx_b,  Syy_ab = sum w by_a by_b,  Sxy_ab = sum w bx_a by_b
// in registers, and appends, for each of its row nodes a, one entry per
// column node b carrying the 2x2 block K_ab: one reservation atomic per
// (element, row node) instead of a list walk and an atomic per scalar.
// Equal columns are merged later by the per-row sort in consolidation.
#include "alls_internal.h"

namespace {

// Gauss abscissa of the reference's 2-point rule, Acc:47 (same literal).
constexpr double GP2 = 0.577350269189626;

__device__ __forceinline__ void store_entry(const ListPool& p, int64_t idx, int32_t col, int32_t eid,
                                            double k00, double k01, double k10, double k11) {
    p.keys[idx] = make_int2(col, eid);
    double2* v = reinterpret_cast<double2*>(static_cast<double*>(p.vals) + idx * 4);
    v[0] = make_double2(k00, k01);
    v[1] = make_double2(k10, k11);
}

// MODE 0: all Q4, 1: all T3, 2: mixed (per-element type array).
template <int MODE>
__global__ __launch_bounds__(256) void k_insert_2d(DevMesh m, ListPool p) {
    const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= m.n_elems) return;
    const int kind = MODE == 2 ? (int)m.types[e] : MODE;
    const int nn = kind == 0 ? 4 : 3;
    if (kind > 1) {
        st_agent(&p.misc[MISC_ERR + ERR_INVALID_MESH], 1);
        return;
    }

    int32_t nd[4];
    const int32_t* c = m.conn + e * m.npe;
    if (MODE != 1 && m.npe == 4) {
        int4 q = *reinterpret_cast<const int4*>(c);
        nd[0] = q.x; nd[1] = q.y; nd[2] = q.z; nd[3] = q.w;
    } else {
        nd[0] = c[0]; nd[1] = c[1]; nd[2] = c[2]; nd[3] = nn == 4 ? c[3] : 0;
    }
    bool ok = true;
    for (int i = 0; i < nn; ++i) ok &= (nd[i] >= 0) & (nd[i] < m.n_nodes);
    if (!ok) {
        st_agent(&p.misc[MISC_ERR + ERR_INVALID_MESH], 1);
        return;
    }
    double x[4], y[4];
    for (int i = 0; i < nn; ++i) {
        double2 q = *reinterpret_cast<const double2*>(m.coords + 2 * (int64_t)nd[i]);
        x[i] = q.x;
        y[i] = q.y;
    }
    const double E = m.E_elem ? m.E_elem[e] : m.E;
    const double nu = m.nu;
    const double f = E / (1.0 - nu * nu);        // D = f [[1,nu,0],[nu,1,0],[0,0,(1-nu)/2]]
    const double cs = (1.0 - nu) / 2.0;

    // Moments over the Gauss points (upper triangles of Sxx, Syy; full Sxy).
    double Sxx[4][4], Syy[4][4], Sxy[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Sxx[a][b] = Syy[a][b] = Sxy[a][b] = 0.0;

    if (kind == 0) {
        // 2x2 Gauss, weights 1 (Acc:47-48); detJ clamped at 1e-10 (Acc:143).
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const double xi = (g & 2) ? GP2 : -GP2;
            const double eta = (g & 1) ? GP2 : -GP2;
            const double dxi[4] = {-0.25 * (1.0 - eta), 0.25 * (1.0 - eta), 0.25 * (1.0 + eta), -0.25 * (1.0 + eta)};
            const double deta[4] = {-0.25 * (1.0 - xi), -0.25 * (1.0 + xi), 0.25 * (1.0 + xi), 0.25 * (1.0 - xi)};
            double J0 = 0, J1 = 0, J2 = 0, J3 = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                J0 += dxi[i] * x[i];
                J1 += dxi[i] * y[i];
                J2 += deta[i] * x[i];
                J3 += deta[i] * y[i];
            }
            double det = J0 * J3 - J1 * J2;
            if (fabs(det) < 1e-10) det = 1e-10;
            const double i0 = J3 / det, i1 = -J1 / det, i2 = -J2 / det, i3 = J0 / det;
            double bx[4], by[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                bx[i] = i0 * dxi[i] + i1 * deta[i];
                by[i] = i2 * dxi[i] + i3 * deta[i];
            }
            const double w = det * m.thickness;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double wx = w * bx[a], wy = w * by[a];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (b >= a) {
                        Sxx[a][b] += wx * bx[b];
                        Syy[a][b] += wy * by[b];
                    }
                    Sxy[a][b] += wx * by[b];
                }
            }
        }
    } else {
        // Constant-strain triangle, one point, weight 0.5*detJ*t (T3:177-205).
        const double J0 = x[1] - x[0], J1 = y[1] - y[0], J2 = x[2] - x[0], J3 = y[2] - y[0];
        const double det = J0 * J3 - J1 * J2;
        const double i0 = J3 / det, i1 = -J1 / det, i2 = -J2 / det, i3 = J0 / det;
        const double bx[3] = {-i0 - i1, i0, i1};
        const double by[3] = {-i2 - i3, i2, i3};
        const double w = 0.5 * det * m.thickness;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                if (b >= a) {
                    Sxx[a][b] = w * bx[a] * bx[b];
                    Syy[a][b] = w * by[a] * by[b];
                }
                Sxy[a][b] = w * bx[a] * by[b];
            }
    }

    // Phase 1: reserve nn slots in each of the nn node rows.
    int32_t pos[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nn) pos[i] = list_reserve(p, nd[i], nn);
    // Phase 2: publish the overflow chunks these reservations own.  Every lane
    // of the wave finishes this before any lane waits on a chunk in phase 3.
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nn) list_own_chunks(p, nd[i], pos[i], nn);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // Phase 3: write the entries.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= nn) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= nn) continue;
            const int32_t s = pos[i] + j;
            const int64_t ch = list_chunk(p, nd[i], s);
            if (ch < 0) continue;
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            const double sxx = Sxx[lo][hi], syy = Syy[lo][hi];
            const double sxy = Sxy[i][j], syx = Sxy[j][i];
            store_entry(p, ch * p.C + (s % p.C), nd[j], (int32_t)e,
                        f * (sxx + cs * syy), f * (nu * sxy + cs * syx),
                        f * (nu * syx + cs * sxy), f * (syy + cs * sxx));
        }
    }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += stride) p[k] = v;
}

}  // namespace

namespace alls {

void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_insert_2d(const DevMesh& m, const ListPool& p, int mode, int block, hipStream_t s) {
    if (m.n_elems <= 0) return;
    if (block != 64 && block != 128 && block != 256) block = 256;
    const int64_t blocks = (m.n_elems + block - 1) / block;
    if (mode == 0)
        hipLaunchKernelGGL(k_insert_2d<0>, dim3((unsigned)blocks), dim3(block), 0, s, m, p);
    else if (mode == 1)
        hipLaunchKernelGGL(k_insert_2d<1>, dim3((unsigned)blocks), dim3(block), 0, s, m, p);
    else
        hipLaunchKernelGGL(k_insert_2d<2>, dim3((unsigned)blocks), dim3(block), 0, s, m, p);
}

}  // namespace alls
