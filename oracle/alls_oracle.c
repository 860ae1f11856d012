/*
 * alls_oracle.c -- CPU oracle (test infrastructure only; see alls_oracle.h).
 *
 * A plain-C restatement of the reference's serial assembly path.  Every
 * element routine keeps the reference's loop order and constants so that the
 * rounding matches its host code; the pattern builder produces the same
 * sorted-unique rows as the reference's std::set builder with a sort+unique
 * over node neighbourhoods instead of a std::set per row.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "alls_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Gauss abscissae and weights exactly as the reference writes them:
 * 2-point rule Acc:47-48 (the literal is 1/sqrt(3) truncated to 15 digits),
 * 3-point rule Hex:45-46. */
static const double GP2[2] = { -0.577350269189626, 0.577350269189626 };
static const double GW2[2] = { 1.0, 1.0 };
static const double GP3[3] = { -0.774596669241483, 0.0, 0.774596669241483 };
static const double GW3[3] = { 0.555555555555556, 0.888888888888889, 0.555555555555556 };

static int nodes_of_kind(int kind) {
    return kind == OR_Q4 ? 4 : kind == OR_T3 ? 3 : kind == OR_HEX20 ? 20 : -1;
}

/* Plane-stress constitutive matrix, Acc:168-176 (E per element here). */
static void d_plane_stress(double E, double nu, double* D) {
    double factor = E / (1.0 - nu * nu);
    D[0] = factor;      D[1] = factor * nu; D[2] = 0.0;
    D[3] = factor * nu; D[4] = factor;      D[5] = 0.0;
    D[6] = 0.0;         D[7] = 0.0;         D[8] = factor * (1.0 - nu) / 2.0;
}

/* ---------------------------------------------------------------- Q4 ---- */
/* Shape-function derivatives, Acc:225-240 (host mirror of Acc:111-126). */
static void q4_dshape(double xi, double eta, double* dNdxi, double* dNdeta) {
    dNdxi[0] = -0.25 * (1.0 - eta);
    dNdxi[1] = 0.25 * (1.0 - eta);
    dNdxi[2] = 0.25 * (1.0 + eta);
    dNdxi[3] = -0.25 * (1.0 + eta);
    dNdeta[0] = -0.25 * (1.0 - xi);
    dNdeta[1] = -0.25 * (1.0 + xi);
    dNdeta[2] = 0.25 * (1.0 + xi);
    dNdeta[3] = 0.25 * (1.0 - xi);
}

/* Jacobian, clamped determinant and B (3x8), Acc:242-280. */
static void q4_jacobian_b(const double* dNdxi, const double* dNdeta, const double* x,
                          const double* y, double* detJ_out, double* B) {
    double J[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (int i = 0; i < 4; i++) {
        J[0] += dNdxi[i] * x[i];
        J[1] += dNdxi[i] * y[i];
        J[2] += dNdeta[i] * x[i];
        J[3] += dNdeta[i] * y[i];
    }
    double detJ = J[0] * J[3] - J[1] * J[2];
    if (fabs(detJ) < 1e-10) detJ = 1e-10;             /* clamp, Acc:257 */
    double invJ[4];
    invJ[0] = J[3] / detJ;
    invJ[1] = -J[1] / detJ;
    invJ[2] = -J[2] / detJ;
    invJ[3] = J[0] / detJ;
    for (int i = 0; i < 4; i++) {
        double dNdx = invJ[0] * dNdxi[i] + invJ[1] * dNdeta[i];
        double dNdy = invJ[2] * dNdxi[i] + invJ[3] * dNdeta[i];
        int col = i * 2;
        B[0 * 8 + col] = dNdx;  B[0 * 8 + col + 1] = 0.0;
        B[1 * 8 + col] = 0.0;   B[1 * 8 + col + 1] = dNdy;
        B[2 * 8 + col] = dNdy;  B[2 * 8 + col + 1] = dNdx;
    }
    *detJ_out = detJ;
}

/* Element stiffness, Acc:292-326: sum over 2x2 Gauss points of B^T D B
 * w detJ t with the m,n,k,l loop order of Acc:313-323. */
void or_ke_q4(const double* x, const double* y, double E, double nu, double t,
              double* ke, double* ke_abs) {
    double D[9];
    d_plane_stress(E, nu, D);
    for (int i = 0; i < 64; i++) ke[i] = 0.0;
    if (ke_abs) for (int i = 0; i < 64; i++) ke_abs[i] = 0.0;
    for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) {
            double xi = GP2[i], eta = GP2[j];
            double weight = GW2[i] * GW2[j];
            double dNdxi[4], dNdeta[4], B[24], detJ;
            q4_dshape(xi, eta, dNdxi, dNdeta);
            q4_jacobian_b(dNdxi, dNdeta, x, y, &detJ, B);
            double iw = weight * detJ * t;
            for (int m = 0; m < 8; m++) {
                for (int n = 0; n < 8; n++) {
                    double sum = 0.0, asum = 0.0;
                    for (int k = 0; k < 3; k++) {
                        for (int l = 0; l < 3; l++) {
                            double p = B[k * 8 + m] * D[k * 3 + l] * B[l * 8 + n];
                            sum += p;
                            asum += fabs(p);
                        }
                    }
                    ke[m * 8 + n] += sum * iw;
                    if (ke_abs) ke_abs[m * 8 + n] += asum * fabs(iw);
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- T3 ---- */
/* Constant-strain triangle, T3:177-205 (identical in Mix:193-226): one
 * point, weight 0.5*detJ*t, no determinant clamp, D*B formed first. */
void or_ke_t3(const double* x, const double* y, double E, double nu, double t,
              double* ke, double* ke_abs) {
    static const double dNdxi[3] = { -1.0, 1.0, 0.0 };   /* T3:131-134 */
    static const double dNdeta[3] = { -1.0, 0.0, 1.0 };
    double D[9];
    d_plane_stress(E, nu, D);
    double J[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (int k = 0; k < 3; k++) {
        J[0] += dNdxi[k] * x[k]; J[1] += dNdxi[k] * y[k];
        J[2] += dNdeta[k] * x[k]; J[3] += dNdeta[k] * y[k];
    }
    double detJ = J[0] * J[3] - J[1] * J[2];
    double invJ[4] = { J[3] / detJ, -J[1] / detJ, -J[2] / detJ, J[0] / detJ };
    double iw = 0.5 * detJ * t;
    double B[18];
    for (int i = 0; i < 18; i++) B[i] = 0.0;
    for (int k = 0; k < 3; k++) {
        double dNdx = invJ[0] * dNdxi[k] + invJ[1] * dNdeta[k];
        double dNdy = invJ[2] * dNdxi[k] + invJ[3] * dNdeta[k];
        B[0 * 6 + k * 2] = dNdx; B[1 * 6 + k * 2 + 1] = dNdy;
        B[2 * 6 + k * 2] = dNdy; B[2 * 6 + k * 2 + 1] = dNdx;
    }
    for (int r = 0; r < 6; r++) {
        for (int c = 0; c < 6; c++) {
            double sum = 0.0, asum = 0.0;
            for (int k = 0; k < 3; k++) {
                double d_val = 0.0, a_val = 0.0;
                for (int l = 0; l < 3; l++) {
                    d_val += D[k * 3 + l] * B[l * 6 + c];
                    a_val += fabs(D[k * 3 + l] * B[l * 6 + c]);
                }
                sum += B[k * 6 + r] * d_val;
                asum += fabs(B[k * 6 + r]) * a_val;
            }
            ke[r * 6 + c] = sum * iw;
            if (ke_abs) ke_abs[r * 6 + c] = asum * fabs(iw);
        }
    }
}

/* ------------------------------------------------------------- Hex20 ---- */
/* Natural coordinates of the 20 nodes, Hex:97-99. */
static const double HXI[20]   = {-1, 1, 1, -1, -1, 1, 1, -1,  0, 1, 0, -1,  0, 1, 0, -1, -1, 1, 1, -1};
static const double HETA[20]  = {-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1};
static const double HZETA[20] = {-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, 0, 0, 0, 0};

/* Serendipity shape-function derivatives, Hex:95-128 (FP64). */
static void hex20_dshape(double xi, double eta, double zeta, double* dNdxi,
                         double* dNdeta, double* dNdzeta) {
    for (int i = 0; i < 8; i++) {
        double a = HXI[i], b = HETA[i], c = HZETA[i];
        dNdxi[i] = 0.125 * a * (1.0 + b * eta) * (1.0 + c * zeta) * (2.0 * a * xi + b * eta + c * zeta - 1.0);
        dNdeta[i] = 0.125 * b * (1.0 + a * xi) * (1.0 + c * zeta) * (a * xi + 2.0 * b * eta + c * zeta - 1.0);
        dNdzeta[i] = 0.125 * c * (1.0 + a * xi) * (1.0 + b * eta) * (a * xi + b * eta + 2.0 * c * zeta - 1.0);
    }
    static const int g1[4] = { 8, 10, 12, 14 };        /* xi_i = 0 */
    for (int q = 0; q < 4; q++) {
        int i = g1[q];
        double b = HETA[i], c = HZETA[i];
        dNdxi[i] = -0.5 * xi * (1.0 + b * eta) * (1.0 + c * zeta);
        dNdeta[i] = 0.25 * b * (1.0 - xi * xi) * (1.0 + c * zeta);
        dNdzeta[i] = 0.25 * c * (1.0 - xi * xi) * (1.0 + b * eta);
    }
    static const int g2[4] = { 9, 11, 13, 15 };        /* eta_i = 0 */
    for (int q = 0; q < 4; q++) {
        int i = g2[q];
        double a = HXI[i], c = HZETA[i];
        dNdxi[i] = 0.25 * a * (1.0 - eta * eta) * (1.0 + c * zeta);
        dNdeta[i] = -0.5 * eta * (1.0 + a * xi) * (1.0 + c * zeta);
        dNdzeta[i] = 0.25 * c * (1.0 + a * xi) * (1.0 - eta * eta);
    }
    for (int i = 16; i < 20; i++) {                    /* zeta_i = 0 */
        double a = HXI[i], b = HETA[i];
        dNdxi[i] = 0.25 * a * (1.0 + b * eta) * (1.0 - zeta * zeta);
        dNdeta[i] = 0.25 * b * (1.0 + a * xi) * (1.0 - zeta * zeta);
        dNdzeta[i] = -0.5 * zeta * (1.0 + a * xi) * (1.0 + b * eta);
    }
}

/* Element stiffness, Hex:130-195 promoted to FP64: 3x3x3 Gauss, detJ clamp
 * at 1e-15 (Hex:156), explicit 3x3 block per node pair (Hex:173-191). */
void or_ke_hex20(const double* x, const double* y, const double* z, double E, double nu,
                 double* ke, double* ke_abs) {
    double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1.0 + nu));
    double c1 = lambda + 2.0 * mu, c2 = lambda, c3 = mu;
    for (int i = 0; i < 3600; i++) ke[i] = 0.0;
    if (ke_abs) for (int i = 0; i < 3600; i++) ke_abs[i] = 0.0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++) {
                double xi = GP3[i], eta = GP3[j], zeta = GP3[k];
                double weight = GW3[i] * GW3[j] * GW3[k];
                double dNdxi[20], dNdeta[20], dNdzeta[20];
                hex20_dshape(xi, eta, zeta, dNdxi, dNdeta, dNdzeta);
                double J[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
                for (int m = 0; m < 20; m++) {
                    J[0] += dNdxi[m] * x[m]; J[1] += dNdxi[m] * y[m]; J[2] += dNdxi[m] * z[m];
                    J[3] += dNdeta[m] * x[m]; J[4] += dNdeta[m] * y[m]; J[5] += dNdeta[m] * z[m];
                    J[6] += dNdzeta[m] * x[m]; J[7] += dNdzeta[m] * y[m]; J[8] += dNdzeta[m] * z[m];
                }
                double detJ = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6])
                            + J[2] * (J[3] * J[7] - J[4] * J[6]);
                if (fabs(detJ) < 1e-15) detJ = 1e-15;
                double invJ[9];
                invJ[0] = (J[4] * J[8] - J[5] * J[7]) / detJ; invJ[1] = (J[2] * J[7] - J[1] * J[8]) / detJ;
                invJ[2] = (J[1] * J[5] - J[2] * J[4]) / detJ; invJ[3] = (J[5] * J[6] - J[3] * J[8]) / detJ;
                invJ[4] = (J[0] * J[8] - J[2] * J[6]) / detJ; invJ[5] = (J[2] * J[3] - J[0] * J[5]) / detJ;
                invJ[6] = (J[3] * J[7] - J[4] * J[6]) / detJ; invJ[7] = (J[1] * J[6] - J[0] * J[7]) / detJ;
                invJ[8] = (J[0] * J[4] - J[1] * J[3]) / detJ;
                double dx[20], dy[20], dz[20];
                for (int m = 0; m < 20; m++) {
                    dx[m] = invJ[0] * dNdxi[m] + invJ[1] * dNdeta[m] + invJ[2] * dNdzeta[m];
                    dy[m] = invJ[3] * dNdxi[m] + invJ[4] * dNdeta[m] + invJ[5] * dNdzeta[m];
                    dz[m] = invJ[6] * dNdxi[m] + invJ[7] * dNdeta[m] + invJ[8] * dNdzeta[m];
                }
                double w = weight * detJ, aw = fabs(w);
                for (int I = 0; I < 20; I++) {
                    for (int Jn = 0; Jn < 20; Jn++) {
                        int r0 = I * 3, c0 = Jn * 3;
                        double xI = dx[I], yI = dy[I], zI = dz[I];
                        double xJ = dx[Jn], yJ = dy[Jn], zJ = dz[Jn];
                        ke[(r0 + 0) * 60 + c0 + 0] += (c1 * xI * xJ + c3 * yI * yJ + c3 * zI * zJ) * w;
                        ke[(r0 + 0) * 60 + c0 + 1] += (c2 * xI * yJ + c3 * yI * xJ) * w;
                        ke[(r0 + 0) * 60 + c0 + 2] += (c2 * xI * zJ + c3 * zI * xJ) * w;
                        ke[(r0 + 1) * 60 + c0 + 0] += (c3 * xI * yJ + c2 * yI * xJ) * w;
                        ke[(r0 + 1) * 60 + c0 + 1] += (c3 * xI * xJ + c1 * yI * yJ + c3 * zI * zJ) * w;
                        ke[(r0 + 1) * 60 + c0 + 2] += (c2 * yI * zJ + c3 * zI * yJ) * w;
                        ke[(r0 + 2) * 60 + c0 + 0] += (c3 * xI * zJ + c2 * zI * xJ) * w;
                        ke[(r0 + 2) * 60 + c0 + 1] += (c3 * yI * zJ + c2 * zI * yJ) * w;
                        ke[(r0 + 2) * 60 + c0 + 2] += (c3 * xI * xJ + c3 * yI * yJ + c1 * zI * zJ) * w;
                        if (ke_abs) {
                            double ax = fabs(xI * xJ), ay = fabs(yI * yJ), az = fabs(zI * zJ);
                            double axy = fabs(xI * yJ), ayx = fabs(yI * xJ), axz = fabs(xI * zJ);
                            double azx = fabs(zI * xJ), ayz = fabs(yI * zJ), azy = fabs(zI * yJ);
                            double l = fabs(c2), mm = fabs(c3), c = fabs(c1);
                            ke_abs[(r0 + 0) * 60 + c0 + 0] += (c * ax + mm * ay + mm * az) * aw;
                            ke_abs[(r0 + 0) * 60 + c0 + 1] += (l * axy + mm * ayx) * aw;
                            ke_abs[(r0 + 0) * 60 + c0 + 2] += (l * axz + mm * azx) * aw;
                            ke_abs[(r0 + 1) * 60 + c0 + 0] += (mm * axy + l * ayx) * aw;
                            ke_abs[(r0 + 1) * 60 + c0 + 1] += (mm * ax + c * ay + mm * az) * aw;
                            ke_abs[(r0 + 1) * 60 + c0 + 2] += (l * ayz + mm * azy) * aw;
                            ke_abs[(r0 + 2) * 60 + c0 + 0] += (mm * axz + l * azx) * aw;
                            ke_abs[(r0 + 2) * 60 + c0 + 1] += (mm * ayz + l * azy) * aw;
                            ke_abs[(r0 + 2) * 60 + c0 + 2] += (mm * ax + mm * ay + c * az) * aw;
                        }
                    }
                }
            }
        }
    }
}

/* Hex:95-195 in the reference's own float arithmetic and float literals. */
void or_ke_hex20_f32(const float* x, const float* y, const float* z, float E, float nu, float* ke) {
    static const float gp[3] = { -0.774596669241483f, 0.0f, 0.774596669241483f };
    static const float gw[3] = { 0.555555555555556f, 0.888888888888889f, 0.555555555555556f };
    float lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    float mu = E / (2.0f * (1.0f + nu));
    float c1 = lambda + 2.0f * mu, c2 = lambda, c3 = mu;
    for (int i = 0; i < 3600; i++) ke[i] = 0.0f;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) {
        float xi = gp[i], eta = gp[j], zeta = gp[k];
        float weight = gw[i] * gw[j] * gw[k];
        float dNdxi[20], dNdeta[20], dNdzeta[20];
        for (int n = 0; n < 8; n++) {
            float a = (float)HXI[n], b = (float)HETA[n], c = (float)HZETA[n];
            dNdxi[n] = 0.125f * a * (1.0f + b * eta) * (1.0f + c * zeta) * (2.0f * a * xi + b * eta + c * zeta - 1.0f);
            dNdeta[n] = 0.125f * b * (1.0f + a * xi) * (1.0f + c * zeta) * (a * xi + 2.0f * b * eta + c * zeta - 1.0f);
            dNdzeta[n] = 0.125f * c * (1.0f + a * xi) * (1.0f + b * eta) * (a * xi + b * eta + 2.0f * c * zeta - 1.0f);
        }
        static const int g1[4] = { 8, 10, 12, 14 };
        for (int q = 0; q < 4; q++) {
            int n = g1[q]; float b = (float)HETA[n], c = (float)HZETA[n];
            dNdxi[n] = -0.5f * xi * (1.0f + b * eta) * (1.0f + c * zeta);
            dNdeta[n] = 0.25f * b * (1.0f - xi * xi) * (1.0f + c * zeta);
            dNdzeta[n] = 0.25f * c * (1.0f - xi * xi) * (1.0f + b * eta);
        }
        static const int g2[4] = { 9, 11, 13, 15 };
        for (int q = 0; q < 4; q++) {
            int n = g2[q]; float a = (float)HXI[n], c = (float)HZETA[n];
            dNdxi[n] = 0.25f * a * (1.0f - eta * eta) * (1.0f + c * zeta);
            dNdeta[n] = -0.5f * eta * (1.0f + a * xi) * (1.0f + c * zeta);
            dNdzeta[n] = 0.25f * c * (1.0f + a * xi) * (1.0f - eta * eta);
        }
        for (int n = 16; n < 20; n++) {
            float a = (float)HXI[n], b = (float)HETA[n];
            dNdxi[n] = 0.25f * a * (1.0f + b * eta) * (1.0f - zeta * zeta);
            dNdeta[n] = 0.25f * b * (1.0f + a * xi) * (1.0f - zeta * zeta);
            dNdzeta[n] = -0.5f * zeta * (1.0f + a * xi) * (1.0f + b * eta);
        }
        float J[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int m = 0; m < 20; m++) {
            J[0] += dNdxi[m] * x[m]; J[1] += dNdxi[m] * y[m]; J[2] += dNdxi[m] * z[m];
            J[3] += dNdeta[m] * x[m]; J[4] += dNdeta[m] * y[m]; J[5] += dNdeta[m] * z[m];
            J[6] += dNdzeta[m] * x[m]; J[7] += dNdzeta[m] * y[m]; J[8] += dNdzeta[m] * z[m];
        }
        float detJ = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6])
                   + J[2] * (J[3] * J[7] - J[4] * J[6]);
        if (fabsf(detJ) < 1e-15f) detJ = 1e-15f;
        float iv[9];
        iv[0] = (J[4] * J[8] - J[5] * J[7]) / detJ; iv[1] = (J[2] * J[7] - J[1] * J[8]) / detJ;
        iv[2] = (J[1] * J[5] - J[2] * J[4]) / detJ; iv[3] = (J[5] * J[6] - J[3] * J[8]) / detJ;
        iv[4] = (J[0] * J[8] - J[2] * J[6]) / detJ; iv[5] = (J[2] * J[3] - J[0] * J[5]) / detJ;
        iv[6] = (J[3] * J[7] - J[4] * J[6]) / detJ; iv[7] = (J[1] * J[6] - J[0] * J[7]) / detJ;
        iv[8] = (J[0] * J[4] - J[1] * J[3]) / detJ;
        float dx[20], dy[20], dz[20];
        for (int m = 0; m < 20; m++) {
            dx[m] = iv[0] * dNdxi[m] + iv[1] * dNdeta[m] + iv[2] * dNdzeta[m];
            dy[m] = iv[3] * dNdxi[m] + iv[4] * dNdeta[m] + iv[5] * dNdzeta[m];
            dz[m] = iv[6] * dNdxi[m] + iv[7] * dNdeta[m] + iv[8] * dNdzeta[m];
        }
        float w = weight * detJ;
        for (int I = 0; I < 20; I++) for (int Jn = 0; Jn < 20; Jn++) {
            int r0 = I * 3, c0 = Jn * 3;
            float xI = dx[I], yI = dy[I], zI = dz[I], xJ = dx[Jn], yJ = dy[Jn], zJ = dz[Jn];
            ke[(r0 + 0) * 60 + c0 + 0] += (c1 * xI * xJ + c3 * yI * yJ + c3 * zI * zJ) * w;
            ke[(r0 + 0) * 60 + c0 + 1] += (c2 * xI * yJ + c3 * yI * xJ) * w;
            ke[(r0 + 0) * 60 + c0 + 2] += (c2 * xI * zJ + c3 * zI * xJ) * w;
            ke[(r0 + 1) * 60 + c0 + 0] += (c3 * xI * yJ + c2 * yI * xJ) * w;
            ke[(r0 + 1) * 60 + c0 + 1] += (c3 * xI * xJ + c1 * yI * yJ + c3 * zI * zJ) * w;
            ke[(r0 + 1) * 60 + c0 + 2] += (c2 * yI * zJ + c3 * zI * yJ) * w;
            ke[(r0 + 2) * 60 + c0 + 0] += (c3 * xI * zJ + c2 * zI * xJ) * w;
            ke[(r0 + 2) * 60 + c0 + 1] += (c3 * yI * zJ + c2 * zI * yJ) * w;
            ke[(r0 + 2) * 60 + c0 + 2] += (c3 * xI * xJ + c3 * yI * yJ + c1 * zI * zJ) * w;
        }
    }
}

/* ------------------------------------------------------- mesh helpers ---- */
static int elem_kind(const or_mesh* m, int64_t e) {
    return m->types ? (int)m->types[e] : (int)m->uniform_type;
}

static double elem_E(const or_mesh* m, int64_t e) {
    return m->E_elem ? m->E_elem[e] : m->E;
}

/* Ke and DOF map of element e (DOF map: Acc:346-351, T3:240-246, Hex:218-224). */
int or_element_ke(const or_mesh* m, int64_t e, double* ke, double* ke_abs, int32_t* dofs) {
    int kind = elem_kind(m, e);
    int nn = nodes_of_kind(kind);
    const int32_t* c = m->conn + e * m->npe;
    int dpn = m->dim;
    if (kind == OR_HEX20) {
        double x[20], y[20], z[20];
        for (int i = 0; i < 20; i++) {
            x[i] = m->coords[(int64_t)c[i] * 3 + 0];
            y[i] = m->coords[(int64_t)c[i] * 3 + 1];
            z[i] = m->coords[(int64_t)c[i] * 3 + 2];
        }
        or_ke_hex20(x, y, z, elem_E(m, e), m->nu, ke, ke_abs);
    } else {
        double x[4], y[4];
        for (int i = 0; i < nn; i++) {
            x[i] = m->coords[(int64_t)c[i] * 2 + 0];
            y[i] = m->coords[(int64_t)c[i] * 2 + 1];
        }
        if (kind == OR_Q4) or_ke_q4(x, y, elem_E(m, e), m->nu, m->thickness, ke, ke_abs);
        else or_ke_t3(x, y, elem_E(m, e), m->nu, m->thickness, ke, ke_abs);
    }
    for (int i = 0; i < nn; i++)
        for (int d = 0; d < dpn; d++) dofs[i * dpn + d] = c[i] * dpn + d;
    return nn * dpn;
}

static int mesh_valid(const or_mesh* m) {
    if (m->dim != 2 && m->dim != 3) return 0;
    for (int64_t e = 0; e < m->n_elems; e++) {
        int kind = elem_kind(m, e);
        int nn = nodes_of_kind(kind);
        if (nn < 0 || nn > m->npe) return 0;
        if ((kind == OR_HEX20) != (m->dim == 3)) return 0;
        for (int i = 0; i < nn; i++) {
            int32_t n = m->conn[e * m->npe + i];
            if (n < 0 || n >= m->n_nodes) return 0;
        }
    }
    return 1;
}

/* Node -> incident elements (CSR). */
static void incidence(const or_mesh* m, int64_t** off_out, int64_t** list_out) {
    int64_t* off = (int64_t*)calloc((size_t)m->n_nodes + 1, sizeof(int64_t));
    for (int64_t e = 0; e < m->n_elems; e++) {
        int nn = nodes_of_kind(elem_kind(m, e));
        for (int i = 0; i < nn; i++) off[m->conn[e * m->npe + i] + 1]++;
    }
    for (int64_t a = 0; a < m->n_nodes; a++) off[a + 1] += off[a];
    int64_t* fill = (int64_t*)malloc(((size_t)m->n_nodes + 1) * sizeof(int64_t));
    memcpy(fill, off, ((size_t)m->n_nodes + 1) * sizeof(int64_t));
    int64_t* list = (int64_t*)malloc(((size_t)off[m->n_nodes] + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < m->n_elems; e++) {
        int nn = nodes_of_kind(elem_kind(m, e));
        for (int i = 0; i < nn; i++) list[fill[m->conn[e * m->npe + i]]++] = e;
    }
    free(fill);
    *off_out = off;
    *list_out = list;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* Sorted unique neighbour nodes of node a (including a); returns count. */
static int node_neighbours(const or_mesh* m, const int64_t* off, const int64_t* list, int64_t a,
                           int32_t** buf, int* cap) {
    int n = 0;
    for (int64_t k = off[a]; k < off[a + 1]; k++) {
        int64_t e = list[k];
        int nn = nodes_of_kind(elem_kind(m, e));
        if (n + nn > *cap) {
            *cap = 2 * (n + nn);
            *buf = (int32_t*)realloc(*buf, (size_t)*cap * sizeof(int32_t));
        }
        for (int i = 0; i < nn; i++) (*buf)[n++] = m->conn[e * m->npe + i];
    }
    qsort(*buf, (size_t)n, sizeof(int32_t), cmp_i32);
    int u = 0;
    for (int i = 0; i < n; i++)
        if (u == 0 || (*buf)[u - 1] != (*buf)[i]) (*buf)[u++] = (*buf)[i];
    return u;
}

/* Pattern: row (a,d) holds {b*dpn+dc} for every node b sharing an element
 * with a -- the same set the reference inserts into std::set per row
 * (Acc:621-631), emitted in ascending order as std::set iteration does. */
int64_t or_pattern(const or_mesh* m, int64_t* row_ptr, int32_t* col_ind) {
    if (!mesh_valid(m)) return -1;
    int dpn = m->dim;
    int64_t *off, *list;
    incidence(m, &off, &list);
    int cap = 256;
    int32_t* buf = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
    row_ptr[0] = 0;
    for (int64_t a = 0; a < m->n_nodes; a++) {
        int u = node_neighbours(m, off, list, a, &buf, &cap);
        for (int d = 0; d < dpn; d++) {
            int64_t r = a * dpn + d;
            row_ptr[r + 1] = row_ptr[r] + (int64_t)u * dpn;
            if (col_ind) {
                int64_t p = row_ptr[r];
                for (int i = 0; i < u; i++)
                    for (int dc = 0; dc < dpn; dc++) col_ind[p++] = buf[i] * dpn + dc;
            }
        }
    }
    free(buf);
    free(off);
    free(list);
    return row_ptr[m->n_nodes * dpn];
}

int64_t or_pattern_bruteforce(const or_mesh* m, int64_t* row_ptr, int32_t* col_ind) {
    if (!mesh_valid(m)) return -1;
    int64_t nd = m->n_nodes * m->dim;
    if (nd > 8192) return -1;
    unsigned char* bits = (unsigned char*)calloc((size_t)(nd * nd), 1);
    int32_t dofs[60];
    double ke[3600];
    for (int64_t e = 0; e < m->n_elems; e++) {
        int ed = or_element_ke(m, e, ke, NULL, dofs);
        for (int i = 0; i < ed; i++)
            for (int j = 0; j < ed; j++) bits[(int64_t)dofs[i] * nd + dofs[j]] = 1;
    }
    row_ptr[0] = 0;
    for (int64_t r = 0; r < nd; r++) {
        int64_t p = row_ptr[r];
        for (int64_t c = 0; c < nd; c++)
            if (bits[r * nd + c]) {
                if (col_ind) col_ind[p] = (int32_t)c;
                p++;
            }
        row_ptr[r + 1] = p;
    }
    free(bits);
    return row_ptr[nd];
}

/* Serial assembly in element order, Acc:329-380: binary search of each
 * column in its CSR row (Acc:355-378) and "+=" of the Ke entry. */
int64_t or_assemble(const or_mesh* m, const int64_t* row_ptr, const int32_t* col_ind,
                    double* val, double* absval) {
    int64_t nnz = row_ptr[m->n_nodes * m->dim];
    for (int64_t i = 0; i < nnz; i++) val[i] = 0.0;
    if (absval) for (int64_t i = 0; i < nnz; i++) absval[i] = 0.0;
    double* ke = (double*)malloc(3600 * sizeof(double));
    double* ka = absval ? (double*)malloc(3600 * sizeof(double)) : NULL;
    int32_t dofs[60];
    int64_t misses = 0;
    for (int64_t e = 0; e < m->n_elems; e++) {
        int ed = or_element_ke(m, e, ke, ka, dofs);
        for (int i = 0; i < ed; i++) {
            int32_t row = dofs[i];
            int64_t row_start = row_ptr[row];
            int64_t row_end = row_ptr[row + 1] - 1;
            for (int j = 0; j < ed; j++) {
                int32_t col = dofs[j];
                int64_t low = row_start, high = row_end, pos = -1;
                while (low <= high) {
                    int64_t mid = low + (high - low) / 2;
                    if (col_ind[mid] == col) { pos = mid; break; }
                    else if (col_ind[mid] < col) low = mid + 1;
                    else high = mid - 1;
                }
                if (pos >= 0) {
                    val[pos] += ke[i * ed + j];
                    if (absval) absval[pos] += ka[i * ed + j];
                } else {
                    misses++;
                }
            }
        }
    }
    free(ke);
    free(ka);
    return misses;
}

void or_assemble_dense(const or_mesh* m, double* K) {
    int64_t nd = m->n_nodes * m->dim;
    for (int64_t i = 0; i < nd * nd; i++) K[i] = 0.0;
    double* ke = (double*)malloc(3600 * sizeof(double));
    int32_t dofs[60];
    for (int64_t e = 0; e < m->n_elems; e++) {
        int ed = or_element_ke(m, e, ke, NULL, dofs);
        for (int i = 0; i < ed; i++)
            for (int j = 0; j < ed; j++) K[(int64_t)dofs[i] * nd + dofs[j]] += ke[i * ed + j];
    }
    free(ke);
}

/* Acc:896-960: errors over entries with a non-zero reference value; means
 * divided by nnz; Frobenius error relative to the reference norm. */
void or_precision(int64_t nnz, const double* ref, const double* test, double* out5) {
    double ref_norm = 0.0;
    for (int64_t i = 0; i < nnz; i++) ref_norm += ref[i] * ref[i];
    ref_norm = sqrt(ref_norm);
    double max_abs = 0.0, max_rel = 0.0, mean_abs = 0.0, mean_rel = 0.0, fro = 0.0;
    for (int64_t i = 0; i < nnz; i++) {
        if (ref[i] != 0.0) {
            double a = fabs(test[i] - ref[i]);
            double r = a / fabs(ref[i]);
            if (a > max_abs) max_abs = a;
            if (r > max_rel) max_rel = r;
            mean_abs += a;
            mean_rel += r;
            fro += (test[i] - ref[i]) * (test[i] - ref[i]);
        }
    }
    out5[0] = max_abs;
    out5[1] = max_rel;
    out5[2] = nnz ? mean_abs / (double)nnz : 0.0;
    out5[3] = nnz ? mean_rel / (double)nnz : 0.0;
    out5[4] = ref_norm > 0.0 ? sqrt(fro) / ref_norm : sqrt(fro);
}

/* Natural-order first-fit colouring, T3:484-523, with a stamp array in place
 * of the reference's per-element std::vector<bool> (same colours). */
int32_t or_color_greedy(const or_mesh* m, int32_t* colors) {
    int64_t *off, *list;
    incidence(m, &off, &list);
    for (int64_t e = 0; e < m->n_elems; e++) colors[e] = -1;
    int cap = 64;
    int64_t* stamp = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
    for (int i = 0; i < cap; i++) stamp[i] = -1;
    int32_t ncol = 0;
    for (int64_t e = 0; e < m->n_elems; e++) {
        int nn = nodes_of_kind(elem_kind(m, e));
        for (int i = 0; i < nn; i++) {
            int32_t a = m->conn[e * m->npe + i];
            for (int64_t k = off[a]; k < off[a + 1]; k++) {
                int64_t f = list[k];
                if (f != e && colors[f] >= 0) stamp[colors[f]] = e;
            }
        }
        int32_t c = 0;
        while (c < cap && stamp[c] == e) c++;
        if (c >= cap) {
            stamp = (int64_t*)realloc(stamp, (size_t)cap * 2 * sizeof(int64_t));
            for (int i = cap; i < 2 * cap; i++) stamp[i] = -1;
            cap *= 2;
        }
        colors[e] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(stamp);
    free(off);
    free(list);
    return ncol;
}

int64_t or_coloring_conflicts(const or_mesh* m, const int32_t* colors) {
    int64_t *off, *list;
    incidence(m, &off, &list);
    int64_t conflicts = 0;
    for (int64_t a = 0; a < m->n_nodes; a++)
        for (int64_t i = off[a]; i < off[a + 1]; i++)
            for (int64_t j = i + 1; j < off[a + 1]; j++)
                if (list[i] != list[j] && colors[list[i]] == colors[list[j]]) conflicts++;
    free(off);
    free(list);
    return conflicts;
}
