/*
 * alls_oracle.h -- CPU oracle for A-LLS finite-element stiffness assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * serial CPU assembly path (zhou1533/A-LLS-Matrix-Parallel-Assembly).  It is
 * the checker for the HIP library: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links,
 * loads or calls it.
 *
 * Reference file keys used in citations:
 *   Acc = ALLS_Quad2D_Assembly_Accuracy_Analysis.cu
 *   T3  = ALLS_T3_Unstructed_Elements_Assemble_Performance.cu
 *   Mix = ALLS_Q4_T3_Mixed_Elements_Assembly_Performance.cu
 *   Hex = ALLS_Hex20_3D_Assembly_Performance.cu
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   or_ke_q4      closed-form top88 stiffness of the unit square, rectangle
 *                 scaling, symmetry, rigid-body null space, patch energy.
 *   or_ke_t3      closed-form CST stiffness t*A*B^T D B (independent formula),
 *                 null space.
 *   or_ke_hex20   symmetry, 6 rigid-body modes, rank 54, uniform-strain
 *                 energy V*eps^T D eps on affine bricks, Gauss-order
 *                 independence (3^3 vs 4^4 via an independent numpy
 *                 integration).
 *   or_pattern    brute-force dense bitmap on tiny meshes, closed-form nnz.
 *   or_assemble   brute-force dense sum P^T Ke P on tiny meshes; global
 *                 symmetry, rigid-body null space and uniform-strain energy.
 *   or_precision  hand-computed small vectors.
 *   or_color_greedy  validity (no two adjacent elements share a colour).
 * No function is "parity unpinned".
 */
#ifndef ALLS_ORACLE_H
#define ALLS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element kinds; Q4=0 and T3=1 follow the reference's ElementType (T3:37-40). */
enum { OR_Q4 = 0, OR_T3 = 1, OR_HEX20 = 2 };

typedef struct {
    int32_t dim;            /* 2 (Q4/T3/mixed) or 3 (Hex20); also DOF per node */
    int32_t npe;            /* stride of conn (4 for Q4 or mixed, 3 for T3, 20 for Hex20) */
    int64_t n_nodes;
    int64_t n_elems;
    const double*  coords;  /* [n_nodes*dim], interleaved x,y(,z) */
    const int32_t* conn;    /* [n_elems*npe]; unused trailing slots are -1 */
    const uint8_t* types;   /* [n_elems] element kind, or NULL => uniform_type */
    int32_t uniform_type;
    double E;               /* Young's modulus when E_elem == NULL */
    const double* E_elem;   /* per-element Young's modulus or NULL */
    double nu;              /* Poisson ratio */
    double thickness;       /* plane-stress thickness (2D only) */
} or_mesh;

/* Element stiffness matrices (row-major, elem_dof x elem_dof).
 * ke_abs (optional) receives the same sums taken over absolute values of
 * every elementary product: the per-entry rounding scale used by the
 * tolerance |gpu - oracle| <= 1e-12 * sum_e |contributions|. */
void or_ke_q4(const double* x, const double* y, double E, double nu, double t,
              double* ke /*64*/, double* ke_abs /*64 or NULL*/);
void or_ke_t3(const double* x, const double* y, double E, double nu, double t,
              double* ke /*36*/, double* ke_abs /*36 or NULL*/);
void or_ke_hex20(const double* x, const double* y, const double* z, double E, double nu,
                 double* ke /*3600*/, double* ke_abs /*3600 or NULL*/);
/* The reference's own single-precision Hex20 routine (Hex:95-195), kept in
 * float arithmetic to report what the reference's FP32 path would give. */
void or_ke_hex20_f32(const float* x, const float* y, const float* z, float E, float nu,
                     float* ke /*3600*/);

/* Element stiffness of element e of a mesh (dispatches on type); returns elem_dof. */
int or_element_ke(const or_mesh* m, int64_t e, double* ke, double* ke_abs, int32_t* dofs);

/* CSR pattern: per-row sorted unique columns (Acc:605-645, T3:458-482,
 * Mix:488-512, Hex:364-392).  row_ptr has n_nodes*dim+1 entries.  When
 * col_ind is NULL only row_ptr is filled.  Returns nnz, or -1 on invalid mesh. */
int64_t or_pattern(const or_mesh* m, int64_t* row_ptr, int32_t* col_ind);
/* Same output from a dense bitmap (tiny meshes, n_dof <= 8192). */
int64_t or_pattern_bruteforce(const or_mesh* m, int64_t* row_ptr, int32_t* col_ind);

/* Serial element-order assembly into a given pattern (Acc:329-380).
 * absval (optional) receives sum_e |Ke| per entry.  Returns the number of
 * contributions whose column was missing from the pattern (0 expected). */
int64_t or_assemble(const or_mesh* m, const int64_t* row_ptr, const int32_t* col_ind,
                    double* val, double* absval);
/* Dense sum of P^T Ke P (tiny meshes): K is n_dof x n_dof row-major. */
void or_assemble_dense(const or_mesh* m, double* K);

/* Precision metrics of Acc:896-960: out = {max_abs, max_rel, mean_abs,
 * mean_rel, frobenius_rel}. */
void or_precision(int64_t nnz, const double* ref, const double* test, double* out5);

/* Natural-order first-fit element colouring through shared nodes
 * (T3:484-523).  Returns the number of colours. */
int32_t or_color_greedy(const or_mesh* m, int32_t* colors);
/* Number of adjacent element pairs (sharing a node) with equal colour. */
int64_t or_coloring_conflicts(const or_mesh* m, const int32_t* colors);

#ifdef __cplusplus
}
#endif
#endif
