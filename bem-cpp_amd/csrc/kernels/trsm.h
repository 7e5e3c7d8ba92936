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

// Value of lane l (uniform) of the wave, broadcast through the scalar unit.
__device__ __forceinline__ double wave_bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float wave_bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Solves T X = B by substitution for one 256 x 256 triangular block T (unit
// lower: forward; non-unit upper: backward) and the columns of B, CW columns
// per block (blockIdx.x; the column count a multiple of CW).  This is the
// backward-stable form of the FP64 solves and of U12 = L11^{-1} A12 in the
// FP64 factorisation (no explicit inverse is formed).
//   Thread t keeps row t of X (CW columns) and, slab by slab, its own row of
// T (S columns) in registers; the next slab's row is loaded while the
// current one is used, so the triangle never passes through LDS.  Inside a
// slab the S rows are solved by the one wave that holds them: the finished
// row of each step goes to every lane of the wave through the scalar unit
// (readlane), with no LDS round trip and no barrier.  Then the slab's rows go
// to LDS, one barrier publishes them (the LDS rows are double buffered by slab
// parity, so that is the only barrier of a slab) and every row outside the
// slab subtracts their contribution.  The upper case scales the pivot row by
// the reciprocal of its diagonal entry, formed by every lane of the slab for
// its own row before the steps (off the critical path of the steps).  B and X
// may be the same array (each block reads its columns before it writes them).
template <typename T, int CW, int S, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm_rhs(const T* __restrict__ Tr, const T* __restrict__ Ti, long ldt,
                                                  const T* Br, const T* Bi, long ldb, T* Xr, T* Xi, long ldx) {
    __builtin_amdgcn_s_setprio(3);  // U12 of the next panel is on the factorisation's critical path
    constexpr int R = 256, NS = R / S;
    static_assert(S <= 64 && 64 % S == 0, "a slab lies inside one wave");
    __shared__ T pr[2][S][CW], pi[2][S][CW];
    const int t = threadIdx.x, c0 = blockIdx.x * CW;
    const T* rowr = Tr + (long)t * ldt;
    const T* rowi = Ti + (long)t * ldt;
    T xr[CW], xi[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) {
        xr[c] = Br[(long)t * ldb + c0 + c];
        xi[c] = Bi[(long)t * ldb + c0 + c];
    }
    auto slab0 = [](int sI) { return LOWER ? sI * S : R - S * (sI + 1); };
    T ar[S], ai[S], nr[S], ni[S];
#pragma unroll
    for (int q = 0; q < S; q++) {
        ar[q] = rowr[slab0(0) + q];
        ai[q] = rowi[slab0(0) + q];
    }
#pragma unroll 1
    for (int sI = 0; sI < NS; sI++) {
        const int s = slab0(sI), buf = sI & 1;
        if (sI + 1 < NS) {
            const int sn = slab0(sI + 1);
#pragma unroll
            for (int q = 0; q < S; q++) {
                nr[q] = rowr[sn + q];
                ni[q] = rowi[sn + q];
            }
        }
        if ((t >> 6) == (s >> 6)) {  // the wave that holds rows s .. s + S - 1
            T rr = T(1), ri = T(0);   // upper: 1 / T[t][t] for the rows of the slab
            if (!LOWER) {
                T dr = T(1), di = T(0);
#pragma unroll
                for (int q = 0; q < S; q++)
                    if (t == s + q) {
                        dr = ar[q];
                        di = ai[q];
                    }
                const T inv = T(1) / (dr * dr + di * di);
                rr = dr * inv;
                ri = -di * inv;
            }
#pragma unroll
            for (int k = 0; k < S; k++) {
                const int q = LOWER ? k : S - 1 - k, l = s + q;
                if (!LOWER && t == l) {
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        const T a = xr[c], b = xi[c];
                        xr[c] = a * rr - b * ri;
                        xi[c] = a * ri + b * rr;
                    }
                }
                T ur[CW], ui[CW];
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    ur[c] = wave_bcast(xr[c], l & 63);
                    ui[c] = wave_bcast(xi[c], l & 63);
                }
                if (LOWER ? (t > l && t < s + S) : (t >= s && t < l)) {
                    const T a = ar[q], b = ai[q];
                    const T na = -a, nb = -b;
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        xr[c] = fmadd(na, ur[c], xr[c]);
                        xr[c] = fmadd(b, ui[c], xr[c]);
                        xi[c] = fmadd(na, ui[c], xi[c]);
                        xi[c] = fmadd(nb, ur[c], xi[c]);
                    }
                }
            }
            if (t >= s && t < s + S) {
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    pr[buf][t - s][c] = xr[c];
                    pi[buf][t - s][c] = xi[c];
                }
            }
        }
        __syncthreads();  // pr[buf] holds the slab's S solved rows
        if (LOWER ? (t >= s + S) : (t < s)) {
#pragma unroll
            for (int q = 0; q < S; q++) {
                const T a = ar[q], b = ai[q];
                const T na = -a, nb = -b;
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    const T ur = pr[buf][q][c], ui = pi[buf][q][c];
                    xr[c] = fmadd(na, ur, xr[c]);
                    xr[c] = fmadd(b, ui, xr[c]);
                    xi[c] = fmadd(na, ui, xi[c]);
                    xi[c] = fmadd(nb, ur, xi[c]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < S; q++) {
            ar[q] = nr[q];
            ai[q] = ni[q];
        }
    }
#pragma unroll
    for (int c = 0; c < CW; c++) {
        Xr[(long)t * ldx + c0 + c] = xr[c];
        Xi[(long)t * ldx + c0 + c] = xi[c];
    }
}

// T X = B for the 256 x 256 triangle at (Tr, Ti) and ncols columns (a
// multiple of 8) on stream s.  Few columns (the solves' right-hand sides, the
// 256 columns of the next panel): 2 columns and 16-row slabs per block, so the
// latency-bound chain of steps is as short as it gets; many columns (the rest
// of a panel's U12): 8 columns and 8-row slabs, where throughput counts
// (measured on MI355X, FP64: 48 / 61 us lower / upper at 128 columns, 274 /
// 322 us at 9216).
template <typename T>
inline void trsm_rhs(hipStream_t s, bool lower, const T* Tr, const T* Ti, long ldt, const T* Br, const T* Bi,
                     long ldb, T* Xr, T* Xi, long ldx, int ncols) {
    if (ncols <= 512) {
        const dim3 grid((unsigned)(ncols / 2));
        if (lower)
            hipLaunchKernelGGL((k_trsm_rhs<T, 2, 16, true>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr, Xi,
                               ldx);
        else
            hipLaunchKernelGGL((k_trsm_rhs<T, 2, 16, false>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr, Xi,
                               ldx);
    } else {
        const dim3 grid((unsigned)(ncols / 8));
        if (lower)
            hipLaunchKernelGGL((k_trsm_rhs<T, 8, 8, true>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr, Xi,
                               ldx);
        else
            hipLaunchKernelGGL((k_trsm_rhs<T, 8, 8, false>), grid, dim3(256), 0, s, Tr, Ti, ldt, Br, Bi, ldb, Xr, Xi,
                               ldx);
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
