// lu.hip -- blocked complex LU with row pivoting on gfx950, FP64 and FP32.
//
// Replaces cuSOLVER Zgetrf (solver.cu:7-48, 87-113).  The reference factors
// the column-major view of its row-major Z (i.e. Z^T) and solves with OP_T;
// this library factors P Z = L U directly on the row-major matrix (rows are
// contiguous, so row interchanges are contiguous copies) and solves Z X = B.
// The pivots differ from cuSOLVER's; the solution does not.
//
// Right-looking two-level blocking on the padded matrix (npad = 256 k):
//   outer panel of NB = 256 columns, factored in 8 inner steps of PB = 32:
//     1. pivot selection for the 32 columns over all remaining rows by a
//        tournament (communication-avoiding "TSLU": Grigori, Demmel & Xiang,
//        SIAM J. Matrix Anal. Appl. 32 (2011)), in one launch: 256-row chunks
//        each choose 32 candidate rows by partial pivoting in LDS; the last
//        block of every group of 8 to finish (integer arrival counter) merges
//        the group's candidates, up to the root, whose pivot order becomes
//        LAPACK-style swaps ipiv[j0 .. j0+31].  No block ever waits.
//     2. the swaps are applied to whole rows (all columns)
//     3. the 32 x 32 diagonal block is factored without further pivoting,
//        L21 = A21 U11^{-1} row by row, U12 of the rest of the outer panel
//     4. rank-32 update of the rest of the outer panel on the matrix cores
//   then L11^{-1} of the 256 x 256 diagonal block (trsm.h), U12 = L11^{-1} A12
//   and the rank-256 trailing update A22 -= L21 U12 as matrix-core GEMMs
//   (gemm.h; 3M form in FP32).  The inverses of all L and U diagonal blocks
//   are kept for the GEMM-based multi-RHS solves.
#include "../bem_internal.h"
#include "gemm.h"
#include "trsm.h"

namespace {

constexpr int NB = 256;     // outer panel width = solve block size
constexpr int PB = 32;      // inner panel width (pivot block)
constexpr int CH = 256;     // rows per tournament block
constexpr int G = CH / PB;  // candidate sets merged per block

template <typename T>
__device__ __forceinline__ T mag2(T r, T i) {
    return r * r + i * i;
}

// Block-wide argmax of (val, idx) over 256 threads; ties -> smaller idx.
template <typename T>
__device__ __forceinline__ void argmax256(T val, int idx, T* sv, int* si, T& best, int& bidx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(val, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > val || (ov == val && oi < idx)) {
            val = ov;
            idx = oi;
        }
    }
    if (lane == 0) {
        sv[wave] = val;
        si[wave] = idx;
    }
    __syncthreads();
    best = sv[0];
    bidx = si[0];
    for (int w = 1; w < 4; w++)
        if (sv[w] > best || (sv[w] == best && si[w] < bidx)) {
            best = sv[w];
            bidx = si[w];
        }
    __syncthreads();
}

template <typename T>
struct TsluShared {
    T vr[CH][PB + 1], vi[CH][PB + 1];
    int id[CH], src[CH];
    T sv[4];
    int si[4];
    int last;
};

// Partial pivoting on the nrows rows in LDS: rows 0..31 end up as the chosen
// rows in pivot order (id/src permuted along).
template <typename T>
__device__ void gepp_select(TsluShared<T>& S, int nrows) {
    const int t = threadIdx.x;
    for (int jc = 0; jc < PB; jc++) {
        const T v = (t >= jc && t < nrows) ? mag2(S.vr[t][jc], S.vi[t][jc]) : T(-1);
        T best;
        int p;
        argmax256(v, t, S.sv, S.si, best, p);
        if (p != jc && t < PB) {
            T a = S.vr[jc][t];
            S.vr[jc][t] = S.vr[p][t];
            S.vr[p][t] = a;
            a = S.vi[jc][t];
            S.vi[jc][t] = S.vi[p][t];
            S.vi[p][t] = a;
        }
        if (p != jc && t == 0) {
            int x = S.id[jc];
            S.id[jc] = S.id[p];
            S.id[p] = x;
            x = S.src[jc];
            S.src[jc] = S.src[p];
            S.src[p] = x;
        }
        __syncthreads();
        if (best > T(0) && t > jc && t < nrows) {
            T lr, li;
            cdiv(S.vr[t][jc], S.vi[t][jc], S.vr[jc][jc], S.vi[jc][jc], lr, li);
            for (int c = jc + 1; c < PB; c++) {
                const T ur = S.vr[jc][c], ui = S.vi[jc][c];
                S.vr[t][c] -= lr * ur - li * ui;
                S.vi[t][c] -= lr * ui + li * ur;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Pivot tournament for columns j0..j0+31 over rows j0..n-1, one launch.
// Candidate sets of level L live at set index off[L] + i; set = 32 ids +
// 32 x 32 original values per plane.  cnt: one arrival counter per group per
// level (zero on entry; reset by the merging block).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_tslu_tree(const T* __restrict__ Ar, const T* __restrict__ Ai, long ld,
                                                   int j0, int n, int* __restrict__ cid, T* __restrict__ cvr,
                                                   T* __restrict__ cvi, int* __restrict__ cnt, int* __restrict__ ipiv) {
    __shared__ TsluShared<T> S;
    const int t = threadIdx.x;
    // level 0: rows j0 + b*CH .. (coalesced: 32 consecutive columns of a row)
    const int r0 = j0 + blockIdx.x * CH;
    int nrows = min(CH, n - r0);
    for (int e = t; e < nrows * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        const long o = (long)(r0 + r) * ld + j0 + c;
        S.vr[r][c] = Ar[o];
        S.vi[r][c] = Ai[o];
    }
    if (t < nrows) {
        S.id[t] = r0 + t;
        S.src[t] = r0 + t;
    }
    __syncthreads();
    gepp_select(S, nrows);
    // write candidate set = original rows of the chosen ids
    int level_sets = (n - j0 + CH - 1) / CH;  // sets at this level
    int set = blockIdx.x, off = 0, cnt_off = 0;
    for (int e = t; e < PB * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        const long g = (long)S.src[r] * ld + j0 + c;
        cvr[((long)(off + set) * PB + r) * PB + c] = Ar[g];
        cvi[((long)(off + set) * PB + r) * PB + c] = Ai[g];
    }
    if (t < PB) cid[(off + set) * PB + t] = S.id[t];
    while (level_sets > 1) {
        // arrive at the group counter; the last arrival merges the group
        const int grp = set / G;
        const int gsize = min(G, level_sets - grp * G);
        __threadfence();
        __syncthreads();
        if (t == 0) {
            const int prev = atomicAdd(&cnt[cnt_off + grp], 1);
            S.last = prev == gsize - 1;
            if (S.last) cnt[cnt_off + grp] = 0;  // ready for the next launch
        }
        __syncthreads();
        if (!S.last) return;
        __threadfence();
        // load the group's candidate sets (original values)
        const int s0 = off + grp * G;
        nrows = gsize * PB;
        for (int e = t; e < nrows * PB; e += 256) {
            const int r = e / PB, c = e % PB;
            S.vr[r][c] = cvr[((long)s0 * PB + r) * PB + c];
            S.vi[r][c] = cvi[((long)s0 * PB + r) * PB + c];
        }
        if (t < nrows) {
            S.id[t] = cid[s0 * PB + t];
            S.src[t] = s0 * PB + t;
        }
        __syncthreads();
        gepp_select(S, nrows);
        const int noff = off + level_sets;
        cnt_off += (level_sets + G - 1) / G;
        set = grp;
        off = noff;
        level_sets = (level_sets + G - 1) / G;
        for (int e = t; e < PB * PB; e += 256) {
            const int r = e / PB, c = e % PB;
            const long g = (long)S.src[r] * PB + c;
            cvr[((long)(off + set) * PB + r) * PB + c] = cvr[g];
            cvi[((long)(off + set) * PB + r) * PB + c] = cvi[g];
        }
        if (t < PB) cid[(off + set) * PB + t] = S.id[t];
    }
    // root: pivot order -> sequential swaps; a chosen row sitting at the
    // position being vacated moves to the other side of the swap
    if (t == 0) {
        int pos[PB];
        for (int c = 0; c < PB; c++) pos[c] = S.id[c];
        for (int c = 0; c < PB; c++) {
            const int x = j0 + c, y = pos[c];
            ipiv[x] = y;
            if (y != x)
                for (int c2 = c + 1; c2 < PB; c2++)
                    if (pos[c2] == x) pos[c2] = y;
            pos[c] = x;
        }
    }
}

// Applies the swaps ipiv[r0 .. r0+cnt) to every column (one thread per column).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* __restrict__ Ar, T* __restrict__ Ai, long ld, int ncols, int r0,
                                               int cnt, const int* __restrict__ ipiv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    for (int k = 0; k < cnt; k++) {
        const int x = r0 + k, y = ipiv[x];
        if (y == x) continue;
        const long ox = (long)x * ld + c, oy = (long)y * ld + c;
        const T a = Ar[ox], b = Ai[ox];
        Ar[ox] = Ar[oy];
        Ai[ox] = Ai[oy];
        Ar[oy] = a;
        Ai[oy] = b;
    }
}

// Panel step after the swaps.  Every block factors the 32 x 32 diagonal block
// (rows/cols j0..j0+31) in LDS without pivoting.  Block 0 writes it back,
// computes U12 = L11^{-1} A12 for columns j0+32 .. j0+32+nright and records a
// zero pivot in *info.  Blocks >= 1 compute L21 = A21 U11^{-1} for 256 rows
// each, staged through LDS with coalesced loads and stores.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(T* __restrict__ Ar, T* __restrict__ Ai, long ld, int j0, int nrows,
                                               int nright, int* __restrict__ info) {
    __shared__ T ur[PB][PB + 1], ui[PB][PB + 1];
    __shared__ T lr_s[256][PB + 1], li_s[256][PB + 1];
    const int t = threadIdx.x;
    for (int e = t; e < PB * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        ur[r][c] = Ar[(long)(j0 + r) * ld + j0 + c];
        ui[r][c] = Ai[(long)(j0 + r) * ld + j0 + c];
    }
    const int rb = j0 + PB + (blockIdx.x - 1) * 256;  // first L21 row of this block
    const int nr = blockIdx.x == 0 ? 0 : min(256, j0 + nrows - rb);
    for (int e = t; e < nr * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        lr_s[r][c] = Ar[(long)(rb + r) * ld + j0 + c];
        li_s[r][c] = Ai[(long)(rb + r) * ld + j0 + c];
    }
    __syncthreads();
    for (int k = 0; k < PB; k++) {
        const T pr = ur[k][k], pi = ui[k][k];
        const bool zero = pr == T(0) && pi == T(0);
        if (!zero && t > k && t < PB) {
            T a, b;
            cdiv(ur[t][k], ui[t][k], pr, pi, a, b);
            ur[t][k] = a;
            ui[t][k] = b;
            for (int c = k + 1; c < PB; c++) {
                const T x = ur[k][c], y = ui[k][c];
                ur[t][c] -= a * x - b * y;
                ui[t][c] -= a * y + b * x;
            }
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        for (int e = t; e < PB * PB; e += 256) {
            const int r = e / PB, c = e % PB;
            Ar[(long)(j0 + r) * ld + j0 + c] = ur[r][c];
            Ai[(long)(j0 + r) * ld + j0 + c] = ui[r][c];
        }
        if (t == 0) {
            for (int k = 0; k < PB; k++)
                if (ur[k][k] == T(0) && ui[k][k] == T(0)) {
                    if (*info == 0) *info = j0 + k + 1;
                    break;
                }
        }
        // U12 of the rest of the outer panel: unit-lower forward substitution,
        // one column per thread, the column kept in this thread's row of lr_s
        // (block 0 stages no L21 rows)
        for (int c = t; c < nright; c += 256) {
            const long base = (long)j0 * ld + j0 + PB + c;
            for (int r = 0; r < PB; r++) {
                lr_s[t][r] = Ar[base + (long)r * ld];
                li_s[t][r] = Ai[base + (long)r * ld];
            }
            for (int l = 0; l < PB - 1; l++) {
                const T xr = lr_s[t][l], xi = li_s[t][l];
                for (int r = l + 1; r < PB; r++) {
                    const T a = ur[r][l], b = ui[r][l];
                    lr_s[t][r] -= a * xr - b * xi;
                    li_s[t][r] -= a * xi + b * xr;
                }
            }
            for (int r = 0; r < PB; r++) {
                Ar[base + (long)r * ld] = lr_s[t][r];
                Ai[base + (long)r * ld] = li_s[t][r];
            }
        }
        return;
    }
    // L21 rows: x U11 = a  ->  x_c = (a_c - sum_{l<c} x_l U[l][c]) / U[c][c]
    if (t < nr) {
        for (int c = 0; c < PB; c++) {
            T sr = lr_s[t][c], si = li_s[t][c];
            for (int l = 0; l < c; l++) {
                const T xr = lr_s[t][l], xi = li_s[t][l];
                sr -= xr * ur[l][c] - xi * ui[l][c];
                si -= xr * ui[l][c] + xi * ur[l][c];
            }
            T qr = sr, qi = si;
            if (!(ur[c][c] == T(0) && ui[c][c] == T(0))) cdiv(sr, si, ur[c][c], ui[c][c], qr, qi);
            lr_s[t][c] = qr;
            li_s[t][c] = qi;
        }
    }
    __syncthreads();
    for (int e = t; e < nr * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        Ar[(long)(rb + r) * ld + j0 + c] = lr_s[r][c];
        Ai[(long)(rb + r) * ld + j0 + c] = li_s[r][c];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_rows(int rows, int cols, const T* __restrict__ Sr,
                                                   const T* __restrict__ Si, long lds, T* __restrict__ Dr,
                                                   T* __restrict__ Di, long ldd) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)rows * cols) return;
    const int r = (int)(idx / cols), c = (int)(idx % cols);
    Dr[(long)r * ldd + c] = Sr[(long)r * lds + c];
    Di[(long)r * ldd + c] = Si[(long)r * lds + c];
}

// Blocked factorisation of the padded npad x npad matrix at (Ar, Ai).
// work: candidate sets (ids, values), arrival counters, a 256 x n U12 buffer.
template <typename T>
void factor(bem_ctx* c, T* Ar, T* Ai, T* Lir, T* Lii, T* Uir, T* Uii, int* dinfo, int* cid, T* cvr, T* cvi, int* cnt,
            T* ubr, T* ubi) {
    const int n = c->npad;
    const long ld = n;
    hipStream_t s = c->stream;
    for (int k0 = 0; k0 < n; k0 += NB) {
        for (int j0 = k0; j0 < k0 + NB; j0 += PB) {
            const int m = n - j0;
            const int nsets = (m + CH - 1) / CH;
            int kt = kt_mark(c, KS_TSLU);
            hipLaunchKernelGGL(k_tslu_tree<T>, dim3(nsets), dim3(256), 0, s, Ar, Ai, ld, j0, n, cid, cvr, cvi, cnt,
                               c->ipiv);
            kt_close(c, kt, 0.0, 2.0 * sizeof(T) * (double)m * PB);
            kt = kt_mark(c, KS_LASWP);
            hipLaunchKernelGGL(k_laswp<T>, dim3((n + 255) / 256), dim3(256), 0, s, Ar, Ai, ld, n, j0, PB, c->ipiv);
            kt_close(c, kt, 0.0, 8.0 * sizeof(T) * (double)n * PB);
            const int nright = k0 + NB - (j0 + PB);
            const int nblk = 1 + (m - PB + 255) / 256;
            kt = kt_mark(c, KS_PANEL);
            hipLaunchKernelGGL(k_panel<T>, dim3(nblk), dim3(256), 0, s, Ar, Ai, ld, j0, m, nright, dinfo);
            kt_close(c, kt, 4.0 * (double)m * PB * PB, 4.0 * sizeof(T) * (double)m * PB);
            if (nright > 0) {
                const long o21 = (long)(j0 + PB) * ld + j0, o12 = (long)j0 * ld + j0 + PB;
                const long o22 = (long)(j0 + PB) * ld + j0 + PB;
                cgemm_ctx<T>(c, m - PB, nright, PB, Ar + o21, Ai + o21, ld, Ar + o12, Ai + o12, ld, Ar + o22,
                             Ai + o22, ld, 0);
            }
        }
        // L11^{-1} of this diagonal block (kept for the solves)
        const long o11 = (long)k0 * ld + k0;
        const long io = (long)(k0 / NB) * NB * NB;
        int kt = kt_mark(c, KS_TRSM);
        invert_diag_blocks<T>(s, NB, 1, Ar + o11, Ai + o11, ld, Lir + io, Lii + io, true);
        kt_close(c, kt, 8.0 * NB * NB * NB / 6.0, 4.0 * sizeof(T) * (double)NB * NB);
        const int nr = n - (k0 + NB);
        if (nr > 0) {
            const long o12 = (long)k0 * ld + k0 + NB;
            const long o21 = (long)(k0 + NB) * ld + k0, o22 = (long)(k0 + NB) * ld + k0 + NB;
            // U12 = L11^{-1} A12 into the work buffer, trailing update, U12 back in place
            cgemm_ctx<T>(c, NB, nr, NB, Lir + io, Lii + io, NB, Ar + o12, Ai + o12, ld, ubr, ubi, ld, 1);
            cgemm_ctx<T>(c, nr, nr, NB, Ar + o21, Ai + o21, ld, ubr, ubi, ld, Ar + o22, Ai + o22, ld, 0);
            kt = kt_mark(c, KS_OTHER);
            hipLaunchKernelGGL(k_copy_rows<T>, dim3((unsigned)(((long)NB * nr + 255) / 256)), dim3(256), 0, s, NB,
                               nr, ubr, ubi, ld, Ar + o12, Ai + o12, ld);
            kt_close(c, kt, 0.0, 4.0 * sizeof(T) * (double)NB * nr);
        }
    }
    // U^{-1} of all diagonal blocks in one launch
    const int kt = kt_mark(c, KS_TRSM);
    invert_diag_blocks<T>(s, NB, n / NB, Ar, Ai, ld, Uir, Uii, false);
    kt_close(c, kt, 8.0 * (double)n * NB * NB / 6.0, 4.0 * sizeof(T) * (double)n * NB);
}

template <typename T>
__global__ void k_convert(const double* __restrict__ a, T* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

template <typename T>
int factor_prec(bem_ctx* c, T* Ar, T* Ai, void*& inv, size_t& inv_bytes) {
    const int n = c->npad;
    // diagonal-block inverses: L and U, npad x NB per plane each
    const size_t ib = 4 * (size_t)n * NB * sizeof(T);
    if (inv_bytes < ib) {
        if (inv) (void)hipFree(inv);
        inv = nullptr;
        inv_bytes = 0;
        if (hipMalloc(&inv, ib) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "LU: diagonal-block inverses");
        inv_bytes = ib;
    }
    T* Lir = static_cast<T*>(inv);
    T* Lii = Lir + (size_t)n * NB;
    T* Uir = Lii + (size_t)n * NB;
    T* Uii = Uir + (size_t)n * NB;
    // work: info + counters, candidate ids, candidate values, U12 buffer
    const int nsets0 = (n + CH - 1) / CH;
    const size_t ncnt = 2 * (size_t)nsets0 + 64;
    const size_t bytes = 256 + ncnt * sizeof(int) + 2 * (size_t)nsets0 * PB * sizeof(int) +
                         4 * (size_t)nsets0 * PB * PB * sizeof(T) + 2 * (size_t)NB * n * sizeof(T) + 256;
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    char* p = static_cast<char*>(scr);
    int* dinfo = reinterpret_cast<int*>(p);
    int* cnt = dinfo + 64;
    int* cid = cnt + ncnt;
    T* cvr = reinterpret_cast<T*>(cid + 2 * (size_t)nsets0 * PB);
    T* cvi = cvr + 2 * (size_t)nsets0 * PB * PB;
    T* ubr = cvi + 2 * (size_t)nsets0 * PB * PB;
    T* ubi = ubr + (size_t)NB * n;
    BEM_HIP(c, hipMemsetAsync(dinfo, 0, (64 + ncnt) * sizeof(int), c->stream));
    factor<T>(c, Ar, Ai, Lir, Lii, Uir, Uii, dinfo, cid, cvr, cvi, cnt, ubr, ubi);
    return 0;
}

}  // namespace

// Diagonal-block inverses for FP64 factors loaded from the host (bem_set_lu).
int lu_inverses_launch(bem_ctx* c) {
    const int n = c->npad;
    const size_t ib = 4 * (size_t)n * NB * sizeof(double);
    if (c->inv64_bytes < ib) {
        if (c->inv64) (void)hipFree(c->inv64);
        c->inv64 = nullptr;
        c->inv64_bytes = 0;
        if (hipMalloc(&c->inv64, ib) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "LU: diagonal-block inverses");
        c->inv64_bytes = ib;
    }
    double* Lir = static_cast<double*>(c->inv64);
    double* Lii = Lir + (size_t)n * NB;
    double* Uir = Lii + (size_t)n * NB;
    double* Uii = Uir + (size_t)n * NB;
    invert_diag_blocks<double>(c->stream, NB, n / NB, c->Vr, c->Vi, n, Lir, Lii, true);
    invert_diag_blocks<double>(c->stream, NB, n / NB, c->Vr, c->Vi, n, Uir, Uii, false);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lu_factor_launch(bem_ctx* c, int precision) {
    const int n = c->npad;
    const size_t nn = (size_t)n * n;
    hipStream_t s = c->stream;
    int rc;
    if (precision == BEM_F64) {
        BEM_HIP(c, hipMemcpyAsync(c->Vr, c->Zr, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        BEM_HIP(c, hipMemcpyAsync(c->Vi, c->Zi, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        rc = factor_prec<double>(c, c->Vr, c->Vi, c->inv64, c->inv64_bytes);
    } else {
        if (!c->has_Z32) {
            const size_t blocks = (nn + 255) / 256;
            hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, s, c->Zr, c->Wr, nn);
            hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, s, c->Zi, c->Wi, nn);
        }
        rc = factor_prec<float>(c, c->Wr, c->Wi, c->inv32, c->inv32_bytes);
    }
    if (rc) return rc;
    BEM_HIP(c, hipGetLastError());
    void* scr = nullptr;
    if (int e = ctx_scratch(c, 256, &scr)) return e;
    int info = 0;
    BEM_HIP(c, hipMemcpyAsync(&info, scr, sizeof(int), hipMemcpyDeviceToHost, s));
    // swap sequence -> row permutation (row i of P Z is row perm[i] of Z)
    std::vector<int> ip(n), perm(n);
    BEM_HIP(c, hipMemcpyAsync(ip.data(), c->ipiv, n * sizeof(int), hipMemcpyDeviceToHost, s));
    BEM_HIP(c, hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++)
        if (ip[i] != i) std::swap(perm[i], perm[ip[i]]);
    BEM_HIP(c, hipMemcpy(c->perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return info > c->n ? 0 : info;  // a zero pivot inside the identity padding cannot occur
}
