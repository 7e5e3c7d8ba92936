// trsm.h -- triangular solves of a block of right-hand sides, shared by the
// factorisation and the solve phase.
//
// k_trsm_slab solves T X = B for an R x R triangular block T (R <= 256, a
// multiple of 32; the column count a multiple of CW): unit lower triangular (forward) or non-unit upper
// (backward).  Thread t keeps row t of X (CW columns) in registers; the
// triangle is staged through LDS in 32-column slabs with coalesced loads; the
// finished row of each step is broadcast through a double-buffered LDS row,
// so each step costs one barrier.  A batch of diagonal blocks runs in one
// launch (blockIdx.y), and B == nullptr means B = I, which is how the
// diagonal-block inverses used by the GEMM-based solves are formed.
#pragma once

#include <hip/hip_runtime.h>

__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ void cdiv(T ar, T ai, T br, T bi, T& qr, T& qi) {
    const T d = br * br + bi * bi;
    qr = (ar * br + ai * bi) / d;
    qi = (ai * br - ar * bi) / d;
}

// RMAX: largest R of the launch (sizes the LDS slab; 32 for the 32 x 32 blocks
// of the doubling inverse, so that launch fits beside a running GEMM).
template <typename T, int CW, bool LOWER, int RMAX = 256>
__global__ __launch_bounds__(256) void k_trsm_slab(int R, const T* __restrict__ Tr, const T* __restrict__ Ti,
                                                   long ldt, long bst, const T* __restrict__ Br,
                                                   const T* __restrict__ Bi, long ldb, long bsb, T* __restrict__ Xr,
                                                   T* __restrict__ Xi, long ldx, long bsx, int ncols) {
    __builtin_amdgcn_s_setprio(3);  // on the factorisation's critical path beside the trailing GEMM
    __shared__ T sr[RMAX][33], si[RMAX][33];
    __shared__ T xr_s[2][CW], xi_s[2][CW];
    const int t = threadIdx.x, c0 = blockIdx.x * CW;
    Tr += blockIdx.y * bst;
    Ti += blockIdx.y * bst;
    Xr += blockIdx.y * bsx;
    Xi += blockIdx.y * bsx;
    if (Br) {
        Br += blockIdx.y * bsb;
        Bi += blockIdx.y * bsb;
    }
    // Loads use a clamped row so every lane reads in bounds; rows >= R are
    // zeroed by a select, keeping the register arrays free of branches.
    const int tr = min(t, R - 1);
    const T keep = t < R ? T(1) : T(0);
    T xr[CW], xi[CW];
    if (Br) {
#pragma unroll
        for (int c = 0; c < CW; c++) {
            xr[c] = keep * Br[(long)tr * ldb + c0 + c];
            xi[c] = keep * Bi[(long)tr * ldb + c0 + c];
        }
    } else {
#pragma unroll
        for (int c = 0; c < CW; c++) {
            xr[c] = (t == c0 + c) ? T(1) : T(0);
            xi[c] = T(0);
        }
    }
    const int nslab = R / 32;
#pragma unroll 1
    for (int sI = 0; sI < nslab; sI++) {
        const int s = LOWER ? sI * 32 : R - 32 * (sI + 1);
        __syncthreads();
#pragma unroll 4
        for (int e = t; e < R * 32; e += 256) {  // coalesced: 32 consecutive columns of a row
            const int r = e >> 5, cc = e & 31;
            sr[r][cc] = Tr[(long)r * ldt + s + cc];
            si[r][cc] = Ti[(long)r * ldt + s + cc];
        }
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < 32; k++) {
            const int l = LOWER ? s + k : s + 31 - k;
            const int buf = l & 1;
            if (t == l) {
                if (!LOWER) {
                    const T dr = sr[l][l - s], di = si[l][l - s];
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        T qr, qi;
                        cdiv(xr[c], xi[c], dr, di, qr, qi);
                        xr[c] = qr;
                        xi[c] = qi;
                    }
                }
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    xr_s[buf][c] = xr[c];
                    xi_s[buf][c] = xi[c];
                }
            }
            __syncthreads();
            if (LOWER ? (t > l && t < R) : (t < l)) {
                const T a = sr[t][l - s], b = si[t][l - s];
                const T na = -a, nb = -b;
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    const T ur = xr_s[buf][c], ui = xi_s[buf][c];
                    xr[c] = fmadd(na, ur, xr[c]);
                    xr[c] = fmadd(b, ui, xr[c]);
                    xi[c] = fmadd(na, ui, xi[c]);
                    xi[c] = fmadd(nb, ur, xi[c]);
                }
            }
        }
    }
    if (t < R)
#pragma unroll
        for (int c = 0; c < CW; c++) {
            Xr[(long)t * ldx + c0 + c] = xr[c];
            Xi[(long)t * ldx + c0 + c] = xi[c];
        }
}

// Column-chunk width per precision (registers per thread: 2 CW values).
template <typename T>
struct TrsmCW;
template <>
struct TrsmCW<float> {
    static constexpr int value = 32;
};
template <>
struct TrsmCW<double> {
    static constexpr int value = 16;
};

// Inverts `nblk` R x R diagonal blocks of the factor matrix at (Fr, Fi) (ld ldf,
// block b at offset b*R*(ldf+1)) into Inv (block b at b*R*R, ld R).
template <typename T>
inline void invert_diag_blocks(hipStream_t s, int R, int nblk, const T* Fr, const T* Fi, long ldf, T* Ir, T* Ii,
                               bool lower) {
    constexpr int CW = TrsmCW<T>::value;
    const dim3 grid((R + CW - 1) / CW, nblk);
    const long bst = (long)R * (ldf + 1), bsx = (long)R * R;
    if (lower)
        hipLaunchKernelGGL((k_trsm_slab<T, CW, true>), grid, dim3(256), 0, s, R, Fr, Fi, ldf, bst, nullptr, nullptr,
                           0L, 0L, Ir, Ii, (long)R, bsx, R);
    else
        hipLaunchKernelGGL((k_trsm_slab<T, CW, false>), grid, dim3(256), 0, s, R, Fr, Fi, ldf, bst, nullptr, nullptr,
                           0L, 0L, Ir, Ii, (long)R, bsx, R);
}
