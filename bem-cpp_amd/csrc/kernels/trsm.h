// trsm.h -- triangular solves of a block of right-hand sides, shared by the
// factorisation and the solve phase.
//
// k_trsm_slab solves T X = B for an R x R triangular block T (R <= 256, a
// multiple of 32; the column count a multiple of CW): unit lower triangular (forward) or non-unit upper
// (backward).  Thread t keeps row t of X (CW columns) in registers; the
// triangle is staged through LDS in 32-column slabs with coalesced loads; the
// finished row of each step is broadcast through a double-buffered LDS row,
// so each step costs one barrier (two for the upper triangle, whose row is
// divided by CW lanes at once between them).  A batch of diagonal blocks runs in one
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

// RMAX: largest R of the launch (sizes the LDS slab).
template <typename T, int CW, bool LOWER, int RMAX = 256>
__global__ __launch_bounds__(256) void k_trsm_slab(int R, const T* __restrict__ Tr, const T* __restrict__ Ti,
                                                   long ldt, long bst, const T* __restrict__ Br,
                                                   const T* __restrict__ Bi, long ldb, long bsb, T* __restrict__ Xr,
                                                   T* __restrict__ Xi, long ldx, long bsx, int ncols) {
    __builtin_amdgcn_s_setprio(3);  // on the factorisation's critical path beside the trailing GEMM
    __shared__ T sr[RMAX][33], si[RMAX][33];
    __shared__ T xr_s[2][CW], xi_s[2][CW];
    __shared__ T ur_s[2][CW], ui_s[2][CW];  // upper: the pivot row before its division
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
            if (LOWER) {
                if (t == l) {
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        xr_s[buf][c] = xr[c];
                        xi_s[buf][c] = xi[c];
                    }
                }
                __syncthreads();
            } else {
                // the row goes out undivided, CW lanes divide one entry each
                // (the same division per entry as by the row's own thread),
                // and the row's thread takes the quotients back
                if (t == l) {
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        ur_s[buf][c] = xr[c];
                        ui_s[buf][c] = xi[c];
                    }
                }
                __syncthreads();
                if (t < CW) {
                    T qr, qi;
                    cdiv(ur_s[buf][t], ui_s[buf][t], sr[l][l - s], si[l][l - s], qr, qi);
                    xr_s[buf][t] = qr;
                    xi_s[buf][t] = qi;
                }
                __syncthreads();
                if (t == l) {
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        xr[c] = xr_s[buf][c];
                        xi[c] = xi_s[buf][c];
                    }
                }
            }
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

// Inverses of unit lower triangular 32 x 32 blocks, one wave per block
// (blockIdx.x): the block is staged once in LDS, then lane j < 32 forms
// column j of X = L^{-1} by forward substitution in registers, with no
// barrier inside the substitution (the slab kernel above takes one per row).
// x_i starts at delta_ij and takes the terms k = 0 .. i-1 in that order with
// the slab kernel's four fused multiply-adds per complex term, so the result
// is bit for bit that of k_trsm_slab with B = I.
template <typename T>
__global__ __launch_bounds__(64) void k_inv_lower32(const T* __restrict__ Fr, const T* __restrict__ Fi, long ldf,
                                                    long bsf, T* __restrict__ Xr, T* __restrict__ Xi, long ldx,
                                                    long bsx) {
    __builtin_amdgcn_s_setprio(3);  // on the factorisation's critical path beside the trailing GEMM
    __shared__ T lr[32][33], li[32][33];
    Fr += blockIdx.x * bsf;
    Fi += blockIdx.x * bsf;
    Xr += blockIdx.x * bsx;
    Xi += blockIdx.x * bsx;
    const int t = threadIdx.x;
    for (int e = t; e < 32 * 32; e += 64) {
        const int r = e >> 5, c = e & 31;
        lr[r][c] = Fr[(long)r * ldf + c];
        li[r][c] = Fi[(long)r * ldf + c];
    }
    __syncthreads();
    if (t >= 32) return;
    T xr[32], xi[32];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        T ar = (i == t) ? T(1) : T(0), ai = T(0);
#pragma unroll
        for (int k = 0; k < i; k++) {
            const T a = lr[i][k], b = li[i][k];
            const T na = -a, nb = -b;
            ar = fmadd(na, xr[k], ar);
            ar = fmadd(b, xi[k], ar);
            ai = fmadd(na, xi[k], ai);
            ai = fmadd(nb, xr[k], ai);
        }
        xr[i] = ar;
        xi[i] = ai;
    }
#pragma unroll
    for (int i = 0; i < 32; i++) {
        Xr[(long)i * ldx + t] = xr[i];
        Xi[(long)i * ldx + t] = xi[i];
    }
}

// Solves T X = B by substitution for one 256 x 256 triangular block T (unit
// lower: forward; non-unit upper: backward) and the columns of B, CW columns
// per block (blockIdx.x; ncols a multiple of CW).  This is the backward-
// stable form of the FP64 solves and of U12 = L11^{-1} A12 in the FP64
// factorisation (no explicit inverse is formed).
//   Thread t keeps row t of X in registers.  The triangle is staged in
// 32-column slabs.  Inside a slab the 32 rows are solved by the one wave that
// holds them, with the finished row of each step passed through LDS inside
// the wave (LDS operations of one wave complete in issue order, so the step
// needs no workgroup barrier); then a barrier publishes the slab's 32 rows and
// every row outside the slab subtracts their contribution.  Slabs are 16
// columns wide in FP64 and 32 in FP32 (about 68 KB of LDS either way, so the
// kernel fits on a CU beside a running trailing update).  Three barriers per
// slab instead of one or two per row.  In the upper case the
// pivot row is divided by CW lanes at once, one entry each.  B and X may be
// the same array (each block reads its columns before it writes them).
template <typename T, int CW, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_rhs(const T* __restrict__ Tr, const T* __restrict__ Ti, long ldt,
                                                  const T* Br, const T* Bi, long ldb, T* Xr, T* Xi, long ldx) {
    __builtin_amdgcn_s_setprio(3);  // U12 of the next panel is on the factorisation's critical path
    constexpr int R = 256, S = sizeof(T) == 8 ? 16 : 32;
    __shared__ T sr[R][S + 1], si[R][S + 1];
    __shared__ T pr[S][CW], pi[S][CW];
    const int t = threadIdx.x, c0 = blockIdx.x * CW;
    T xr[CW], xi[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) {
        xr[c] = Br[(long)t * ldb + c0 + c];
        xi[c] = Bi[(long)t * ldb + c0 + c];
    }
#pragma unroll 1
    for (int sI = 0; sI < R / S; sI++) {
        const int s = LOWER ? sI * S : R - S * (sI + 1);
        // rows of this slab's columns that are read: below the slab's top row
        // (lower) or above its bottom row (upper)
        const int r0 = LOWER ? s : 0, r1 = LOWER ? R : s + S;
        __syncthreads();  // the previous slab's readers are done with sr / pr
        for (int e = r0 * S + t; e < r1 * S; e += 256) {
            const int r = e / S, cc = e % S;
            sr[r][cc] = Tr[(long)r * ldt + s + cc];
            si[r][cc] = Ti[(long)r * ldt + s + cc];
        }
        __syncthreads();
        if ((t >> 6) == (s >> 6)) {  // the wave that holds rows s .. s + S - 1
            const int lane0 = t & ~63;
#pragma unroll 1
            for (int k = 0; k < S; k++) {
                const int l = LOWER ? s + k : s + S - 1 - k, q = l - s;
                if (t == l) {
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        pr[q][c] = xr[c];
                        pi[q][c] = xi[c];
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (!LOWER) {
                    if (t - lane0 < CW) {
                        const int c = t - lane0;
                        T qr, qi;
                        cdiv(pr[q][c], pi[q][c], sr[l][q], si[l][q], qr, qi);
                        pr[q][c] = qr;
                        pi[q][c] = qi;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (t == l) {
#pragma unroll
                        for (int c = 0; c < CW; c++) {
                            xr[c] = pr[q][c];
                            xi[c] = pi[q][c];
                        }
                    }
                }
                if (LOWER ? (t > l && t < s + S) : (t >= s && t < l)) {
                    const T a = sr[t][q], b = si[t][q];
                    const T na = -a, nb = -b;
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        const T ur = pr[q][c], ui = pi[q][c];
                        xr[c] = fmadd(na, ur, xr[c]);
                        xr[c] = fmadd(b, ui, xr[c]);
                        xi[c] = fmadd(na, ui, xi[c]);
                        xi[c] = fmadd(nb, ur, xi[c]);
                    }
                }
                // the next step's row goes to another LDS row, but the
                // division of the upper case rewrites this one: keep the
                // reads of this step ahead of the next step's writes
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();  // pr holds the slab's S solved rows
        if (LOWER ? (t >= s + S) : (t < s)) {
#pragma unroll 4
            for (int q = 0; q < S; q++) {
                const T a = sr[t][q], b = si[t][q];
                const T na = -a, nb = -b;
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    const T ur = pr[q][c], ui = pi[q][c];
                    xr[c] = fmadd(na, ur, xr[c]);
                    xr[c] = fmadd(b, ui, xr[c]);
                    xi[c] = fmadd(na, ui, xi[c]);
                    xi[c] = fmadd(nb, ur, xi[c]);
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CW; c++) {
        Xr[(long)t * ldx + c0 + c] = xr[c];
        Xi[(long)t * ldx + c0 + c] = xi[c];
    }
}

// Column-chunk width of k_trsm_rhs (FP64: 16 registers of X per thread).
constexpr int kTrsmRhsCW = 8;

// T X = B for the 256 x 256 triangle at (Tr, Ti) and ncols columns (a
// multiple of kTrsmRhsCW) on stream s.
template <typename T>
inline void trsm_rhs(hipStream_t s, bool lower, const T* Tr, const T* Ti, long ldt, const T* Br, const T* Bi,
                     long ldb, T* Xr, T* Xi, long ldx, int ncols) {
    const dim3 grid((unsigned)(ncols / kTrsmRhsCW));
    if (lower)
        hipLaunchKernelGGL((k_trsm_rhs<T, kTrsmRhsCW, true>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr, Xi,
                           ldx);
    else
        hipLaunchKernelGGL((k_trsm_rhs<T, kTrsmRhsCW, false>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr,
                           Xi, ldx);
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
