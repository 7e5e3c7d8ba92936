// trsm.h -- triangular solves of a block of right-hand sides held in LDS,
// shared by the factorisation (U12 = L11^{-1} A12) and the solve phase.
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
__device__ __forceinline__ void cdiv(T ar, T ai, T br, T bi, T& qr, T& qi) {
    const T d = br * br + bi * bi;
    qr = (ar * br + ai * bi) / d;
    qi = (ai * br - ar * bi) / d;
}

// ---------------------------------------------------------------------------
// Triangular solves with a block of right-hand sides held in LDS (shared with
// the solve phase).  X (R x ncols, row-major, ld ldx) is overwritten by
// L^{-1} X (unit lower) or U^{-1} X (non-unit upper), with L/U the R x R
// diagonal block at (Lr, Li), ld ldl.  One block handles CW columns; thread t
// owns row t (R <= 256).
// ---------------------------------------------------------------------------
template <typename T, int CW>
__global__ __launch_bounds__(256) void k_trsm_lower_unit(int R, const T* __restrict__ Lr, const T* __restrict__ Li,
                                                         long ldl, T* __restrict__ Xr, T* __restrict__ Xi,
                                                         long ldx, int ncols) {
    __shared__ T xr[256][CW + 1], xi[256][CW + 1];
    const int t = threadIdx.x, c0 = blockIdx.x * CW;
    const int nc = min(CW, ncols - c0);
    if (t < R)
        for (int c = 0; c < nc; c++) {
            xr[t][c] = Xr[(long)t * ldx + c0 + c];
            xi[t][c] = Xi[(long)t * ldx + c0 + c];
        }
    __syncthreads();
    for (int l = 0; l < R - 1; l++) {
        if (t > l && t < R) {
            const T a = Lr[(long)t * ldl + l], b = Li[(long)t * ldl + l];
            for (int c = 0; c < nc; c++) {
                xr[t][c] -= a * xr[l][c] - b * xi[l][c];
                xi[t][c] -= a * xi[l][c] + b * xr[l][c];
            }
        }
        __syncthreads();
    }
    if (t < R)
        for (int c = 0; c < nc; c++) {
            Xr[(long)t * ldx + c0 + c] = xr[t][c];
            Xi[(long)t * ldx + c0 + c] = xi[t][c];
        }
}

template <typename T, int CW>
__global__ __launch_bounds__(256) void k_trsm_upper(int R, const T* __restrict__ Ur, const T* __restrict__ Ui,
                                                    long ldu, T* __restrict__ Xr, T* __restrict__ Xi, long ldx,
                                                    int ncols) {
    __shared__ T xr[256][CW + 1], xi[256][CW + 1];
    const int t = threadIdx.x, c0 = blockIdx.x * CW;
    const int nc = min(CW, ncols - c0);
    if (t < R)
        for (int c = 0; c < nc; c++) {
            xr[t][c] = Xr[(long)t * ldx + c0 + c];
            xi[t][c] = Xi[(long)t * ldx + c0 + c];
        }
    __syncthreads();
    for (int l = R - 1; l >= 0; l--) {
        if (t == l) {
            const T dr = Ur[(long)l * ldu + l], di = Ui[(long)l * ldu + l];
            for (int c = 0; c < nc; c++) {
                T qr, qi;
                cdiv(xr[l][c], xi[l][c], dr, di, qr, qi);
                xr[l][c] = qr;
                xi[l][c] = qi;
            }
        }
        __syncthreads();
        if (t < l) {
            const T a = Ur[(long)t * ldu + l], b = Ui[(long)t * ldu + l];
            for (int c = 0; c < nc; c++) {
                xr[t][c] -= a * xr[l][c] - b * xi[l][c];
                xi[t][c] -= a * xi[l][c] + b * xr[l][c];
            }
        }
        __syncthreads();
    }
    if (t < R)
        for (int c = 0; c < nc; c++) {
            Xr[(long)t * ldx + c0 + c] = xr[t][c];
            Xi[(long)t * ldx + c0 + c] = xi[t][c];
        }
}
