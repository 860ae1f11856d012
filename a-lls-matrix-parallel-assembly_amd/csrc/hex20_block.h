// hex20_block.h -- a Hex20 element matrix block from the packed M blocks,
// with its fused multiply-adds written out, so that every kernel that forms
// blocks (list insertion, value-only staging, coloured and atomic scatter)
// does the same operations in the same order and gives the same bits.
//
// M = X^T (W X) is kept as its 210 upper 3x3 node blocks, pair (I <= J) at
// 9 * pair_index(I, J), row-major; block (I, J) of the element matrix is
//   K_IJ = lam * M_IJ + mu * M_IJ^T + mu * tr(M_IJ) * I3
// (M_IJ = M_JI^T for I > J), the sum the reference writes out term by term
// (Hex:173-191).
#pragma once

#include "alls_internal.h"

#ifdef __HIPCC__

__device__ __forceinline__ int pair_index(int I, int J) { return I * (39 - I) / 2 + J; }

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Block (I, J) of K from the packed blocks Mb (LDS or global memory).  The
// fused multiply-adds are written out, so the compiler's contraction choices
// cannot differ between the kernels that use it.
template <typename T>
__device__ __forceinline__ void k_block(const T* Mb, int I, int J, T lam, T mu, T (&kb)[3][3]) {
    const bool up = I <= J;
    const int base = 9 * (up ? pair_index(I, J) : pair_index(J, I));
    T mb[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) mb[r][c] = Mb[base + (up ? 3 * r + c : 3 * c + r)];
    const T mtr = mu * (mb[0][0] + mb[1][1] + mb[2][2]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            kb[r][c] = r == c ? fma_t(lam, mb[r][r], fma_t(mu, mb[r][r], mtr)) : fma_t(lam, mb[r][c], mu * mb[c][r]);
}

#endif  // __HIPCC__
