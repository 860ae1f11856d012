// elem2d.h -- Q4 / T3 plane-stress element stiffness in node-pair blocks,
// shared by the A-LLS insertion kernel and the pattern-scatter kernels.
//
// The reference forms B^T D B per scalar entry with an 8x8 (or 6x6) loop
// over 3x3 products (Acc:199-209, T3:166-170).  Here each element keeps the
// Gauss-point moments
//   Sxx_ab = sum w bx_a bx_b,  Syy_ab = sum w by_a by_b,  Sxy_ab = sum w bx_a by_b
// (bx, by = dN/dx, dN/dy) and forms the 2x2 block of node pair (a,b) from
// them: K_ab = f [[Sxx + c Syy, nu Sxy_ab + c Sxy_ba], [nu Sxy_ba + c Sxy_ab, Syy + c Sxx]]
// with f = E / (1 - nu^2), c = (1 - nu) / 2.  Same sums, about 6x fewer flops.
#pragma once

#include "alls_internal.h"

namespace alls_dev {

// Gauss abscissa of the reference's 2-point rule, Acc:47 (same literal).
constexpr double GP2 = 0.577350269189626;

// Moments of one row node i of an element: the four moment rows that its
// blocks K_ij (j = 0..3) read.  xx/yy[j] = Sxx/Syy of the pair (min(i,j),
// max(i,j)); xy[j] = Sxy_ij; yx[j] = Sxy_ji.
struct RowMoments {
    double xx[4], yy[4], xy[4], yx[4];
};

struct Elem2D {
    int nn;                    // 4 (Q4) or 3 (T3)
    int32_t nd[4];             // node ids
    double E, f, cs, nu, t;
    double x[4], y[4];
    double Sxx[4][4], Syy[4][4], Sxy[4][4];   // Sxx/Syy upper triangles only

    // Loads element e; false (and an error flag) when a node id is out of range.
    // No FMA contraction here: every kernel that integrates an element (list
    // insertion, value-only gather, coloured scatter) then produces the same
    // bits, so the paths agree bitwise, not only within rounding.
    template <int MODE>
    __device__ __forceinline__ bool load(const DevMesh& m, int64_t e, int32_t* misc) {
        if (!load_geom<MODE>(m, e, misc)) return false;
        moments();
        return true;
    }

    // Node ids, coordinates and material of element e, no integration.
    template <int MODE>
    __device__ __forceinline__ bool load_geom(const DevMesh& m, int64_t e, int32_t* misc) {
        if (!load_nodes<MODE>(m, e, misc)) return false;
        load_coords(m, e);
        return true;
    }

    // Node ids of element e (checked against the mesh); false (and an error
    // flag) when a node id is out of range.
    template <int MODE>
    __device__ __forceinline__ bool load_nodes(const DevMesh& m, int64_t e, int32_t* misc) {
        const int kind = MODE == 2 ? (int)m.types[e] : MODE;
        if (kind > 1) {
            st_agent(&misc[MISC_ERR + ERR_INVALID_MESH], 1);
            return false;
        }
        nn = kind == 0 ? 4 : 3;
        const int32_t* c = m.conn + e * m.npe;
        if (MODE != 1 && m.npe == 4) {
            const int4 q = *reinterpret_cast<const int4*>(c);
            nd[0] = q.x; nd[1] = q.y; nd[2] = q.z; nd[3] = q.w;
        } else {
            nd[0] = c[0]; nd[1] = c[1]; nd[2] = c[2]; nd[3] = nn == 4 ? c[3] : 0;
        }
        // Loops run over the 4 slots with a predicate (never up to a runtime
        // bound) so that the arrays stay in registers.
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < nn) ok &= (nd[i] >= 0) & (nd[i] < m.n_nodes);
        if (!ok) {
            st_agent(&misc[MISC_ERR + ERR_INVALID_MESH], 1);
            return false;
        }
        return true;
    }

    // Coordinates and material of element e, after load_nodes.
    __device__ __forceinline__ void load_coords(const DevMesh& m, int64_t e) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x[i] = y[i] = 0.0;
            if (i < nn) {
                const double2 q = *reinterpret_cast<const double2*>(m.coords + 2 * (int64_t)nd[i]);
                x[i] = q.x;
                y[i] = q.y;
            }
        }
        E = m.E_elem ? m.E_elem[e] : m.E;
        nu = m.nu;
        t = m.thickness;
    }

    // f = E / (1 - nu^2) and c = (1 - nu) / 2, after load_geom (kept apart
    // so that a caller can issue memory operations between the loads and the
    // first use of their results).
    __device__ __forceinline__ void material() {
#pragma clang fp contract(off)
        f = E / (1.0 - nu * nu);
        cs = (1.0 - nu) / 2.0;
    }

    // Keeps the moments in registers at this point: the arithmetic that
    // forms them is not moved below the memory operations that follow.
    __device__ __forceinline__ void settle() const {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (b >= a) asm volatile("" ::"v"(Sxx[a][b]), "v"(Syy[a][b]) : "memory");
                asm volatile("" ::"v"(Sxy[a][b]) : "memory");
            }
        asm volatile("" ::"v"(f), "v"(cs) : "memory");
    }

    // Shape-function gradients at Gauss point g (Q4): bx, by and the weight
    // detJ * t.  2x2 Gauss, weights 1 (Acc:47-48); detJ clamped at 1e-10
    // (Acc:143).
    __device__ __forceinline__ double q4_point(int g, double* bx, double* by) const {
#pragma clang fp contract(off)
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
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bx[i] = i0 * dxi[i] + i1 * deta[i];
            by[i] = i2 * dxi[i] + i3 * deta[i];
        }
        return det * t;
    }

    // Constant-strain triangle gradients, weight 0.5*detJ*t, no clamp
    // (T3:177-205).
    __device__ __forceinline__ double t3_point(double* bx, double* by) const {
#pragma clang fp contract(off)
        const double J0 = x[1] - x[0], J1 = y[1] - y[0], J2 = x[2] - x[0], J3 = y[2] - y[0];
        const double det = J0 * J3 - J1 * J2;
        const double i0 = J3 / det, i1 = -J1 / det, i2 = -J2 / det, i3 = J0 / det;
        bx[0] = -i0 - i1; bx[1] = i0; bx[2] = i1; bx[3] = 0.0;
        by[0] = -i2 - i3; by[1] = i2; by[2] = i3; by[3] = 0.0;
        return 0.5 * det * t;
    }

    // Material constants and all moments (Sxx/Syy upper triangles, Sxy full).
    __device__ __forceinline__ void moments() {
#pragma clang fp contract(off)
        material();
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) Sxx[a][b] = Syy[a][b] = Sxy[a][b] = 0.0;
        if (nn == 4) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                double bx[4], by[4];
                const double w = q4_point(g, bx, by);
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
            double bx[4], by[4];
            const double w = t3_point(bx, by);
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
    }

    // The moments of row node i only (i may be a run-time value), with the
    // same operations, in the same order, as moments(): the blocks formed
    // from them carry the same bits.
    __device__ __forceinline__ void row_moments(int i, RowMoments& r) const {
#pragma clang fp contract(off)
#pragma unroll
        for (int b = 0; b < 4; ++b) r.xx[b] = r.yy[b] = r.xy[b] = r.yx[b] = 0.0;
        if (nn == 4) {
#pragma unroll 1
            for (int g = 0; g < 4; ++g) {
                double bx[4], by[4];
                const double w = q4_point(g, bx, by);
                const double bxi = sel4(bx, i), byi = sel4(by, i);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const bool lo = i <= b;   // i is the smaller index of the pair
                    r.xx[b] += (w * (lo ? bxi : bx[b])) * (lo ? bx[b] : bxi);
                    r.yy[b] += (w * (lo ? byi : by[b])) * (lo ? by[b] : byi);
                    r.xy[b] += (w * bxi) * by[b];
                    r.yx[b] += (w * bx[b]) * byi;
                }
            }
        } else {
            double bx[4], by[4];
            const double w = t3_point(bx, by);
            const double bxi = sel4(bx, i), byi = sel4(by, i);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const bool lo = i <= b;
                r.xx[b] = (w * (lo ? bxi : bx[b])) * (lo ? bx[b] : bxi);
                r.yy[b] = (w * (lo ? byi : by[b])) * (lo ? by[b] : byi);
                r.xy[b] = (w * bxi) * by[b];
                r.yx[b] = (w * bx[b]) * byi;
            }
        }
    }

    // 2x2 block K_ij from row node i's moments.
    __device__ __forceinline__ void row_block(const RowMoments& r, int j, double* k) const {
#pragma clang fp contract(off)
        const double sxx = r.xx[j], syy = r.yy[j], sxy = r.xy[j], syx = r.yx[j];
        k[0] = f * (sxx + cs * syy);
        k[1] = f * (nu * sxy + cs * syx);
        k[2] = f * (nu * syx + cs * sxy);
        k[3] = f * (syy + cs * sxx);
    }

    static __device__ __forceinline__ double sel4(const double* v, int i) {
        return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
    }

    // 2x2 block of node pair (i, j): k[0..3] = K00, K01, K10, K11.
    __device__ __forceinline__ void block(int i, int j, double* k) const {
#pragma clang fp contract(off)
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const double sxx = Sxx[lo][hi], syy = Syy[lo][hi];
        const double sxy = Sxy[i][j], syx = Sxy[j][i];
        k[0] = f * (sxx + cs * syy);
        k[1] = f * (nu * sxy + cs * syx);
        k[2] = f * (nu * syx + cs * sxy);
        k[3] = f * (syy + cs * sxx);
    }
};

}  // namespace alls_dev
