/*
 * csr_checks.c -- CSR invariant checks for the full-size parity tests.
 *
 * TEST INFRASTRUCTURE ONLY (built into liballs_oracle.so beside the oracle;
 * loaded by tests/ only).  At the benchmark sizes (Hex20 80^3: nnz 1.09e9)
 * the serial oracle would take minutes, so those outputs are checked by
 * properties that do not depend on size: per-row sorted unique columns,
 * symmetry of the assembled matrix, and products K*u for the rigid-body and
 * uniform-strain fields.  These loops are plain and parallel (OpenMP); they
 * restate no reference code.
 */
#include <math.h>
#include <stdint.h>

/* Number of rows whose columns are not strictly increasing or fall outside
 * [0, n_cols). */
int64_t ck_rows_sorted_unique(int64_t n_rows, const int64_t* rp, const int32_t* ci, int64_t n_cols) {
    int64_t bad = 0;
#pragma omp parallel for schedule(static, 4096) reduction(+ : bad)
    for (int64_t r = 0; r < n_rows; ++r) {
        int ok = rp[r + 1] >= rp[r];
        for (int64_t k = rp[r]; ok && k < rp[r + 1]; ++k) {
            if (ci[k] < 0 || ci[k] >= n_cols) ok = 0;
            if (k > rp[r] && ci[k] <= ci[k - 1]) ok = 0;
        }
        bad += ok ? 0 : 1;
    }
    return bad;
}

/* Position of column c in row r (binary search), or -1. */
static int64_t find_col(const int64_t* rp, const int32_t* ci, int64_t r, int32_t c) {
    int64_t lo = rp[r], hi = rp[r + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (ci[mid] == c) return mid;
        if (ci[mid] < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

/* max over entries of |K_ij - K_ji|; *missing receives the number of entries
 * whose transpose position is absent from the pattern; *max_abs receives
 * max |K_ij|.  Rows must be sorted. */
double ck_symmetry(int64_t n_rows, const int64_t* rp, const int32_t* ci, const double* v, int64_t* missing,
                   double* max_abs) {
    double worst = 0.0, vmax = 0.0;
    int64_t miss = 0;
#pragma omp parallel for schedule(static, 4096) reduction(max : worst, vmax) reduction(+ : miss)
    for (int64_t r = 0; r < n_rows; ++r) {
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            const int32_t c = ci[k];
            const double a = fabs(v[k]);
            if (a > vmax) vmax = a;
            if (c < 0 || c >= n_rows) {
                ++miss;
                continue;
            }
            const int64_t t = find_col(rp, ci, c, (int32_t)r);
            if (t < 0) {
                ++miss;
                continue;
            }
            const double d = fabs(v[k] - v[t]);
            if (d > worst) worst = d;
        }
    }
    if (missing) *missing = miss;
    if (max_abs) *max_abs = vmax;
    return worst;
}

/* y = K x. */
void ck_matvec(int64_t n_rows, const int64_t* rp, const int32_t* ci, const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static, 4096)
    for (int64_t r = 0; r < n_rows; ++r) {
        double s = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
        y[r] = s;
    }
}

/* sum_k |a_k - b_k|^2 and sum_k |b_k|^2 (Frobenius of a - b against b). */
void ck_frobenius(int64_t n, const double* a, const double* b, double* out2) {
    double d2 = 0.0, b2 = 0.0;
#pragma omp parallel for schedule(static, 65536) reduction(+ : d2, b2)
    for (int64_t k = 0; k < n; ++k) {
        const double d = a[k] - b[k];
        d2 += d * d;
        b2 += b[k] * b[k];
    }
    out2[0] = d2;
    out2[1] = b2;
}
