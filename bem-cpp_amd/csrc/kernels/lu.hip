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
//   then L11^{-1} of the 256 x 256 diagonal block (by doubling: eight 32 x 32
//   inverses in LDS, then three levels of batched GEMMs), U12 = L11^{-1} A12
//   and the rank-256 trailing update A22 -= L21 U12 as matrix-core GEMMs
//   (gemm.h; 3M form in FP32).  The inverses of all L and U diagonal blocks
//   are kept for the GEMM-based multi-RHS solves.
#include <cstdlib>
#include <vector>

#include "../bem_internal.h"
#include "gemm.h"
#include "gemm_nt.h"
#include "trsm.h"

namespace {

constexpr int NB = 256;  // outer panel width = solve block size
constexpr int PB = 32;   // inner panel width (pivot block)

// Tournament block: CH rows, one per thread, so a merge takes CH / PB
// candidate sets.  256 threads: larger blocks need fewer tree levels but pay
// more per elimination step (measured slower at 512), and at about 150
// registers the FP32 block still fits beside the trailing-update GEMM on a CU.
template <typename T>
struct Tslu;
template <>
struct Tslu<float> {
    static constexpr int CH = 256;
};
template <>
struct Tslu<double> {
    static constexpr int CH = 256;
};

template <typename T>
__device__ __forceinline__ T mag2(T r, T i) {
    return r * r + i * i;
}

// 16-byte vector type of T (4 floats or 2 doubles).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
};
template <>
struct Vec16<double> {
    using type = double2;
};

// PB consecutive values at p (16-byte aligned) <-> registers.
template <typename T>
__device__ __forceinline__ void load_row(const T* __restrict__ p, T (&x)[PB]) {
    using V = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < PB; q += E) {
        const V a = *reinterpret_cast<const V*>(p + q);
        __builtin_memcpy(&x[q], &a, 16);
    }
}
template <typename T>
__device__ __forceinline__ void store_row(T* __restrict__ p, const T (&x)[PB]) {
    using V = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < PB; q += E) {
        V a;
        __builtin_memcpy(&a, &x[q], 16);
        *reinterpret_cast<V*>(p + q) = a;
    }
}

// 1 / d for the pivot scaling: the hardware reciprocal in FP32 (1 ulp), an
// IEEE division in FP64.
__device__ __forceinline__ float recip(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ double recip(double d) { return 1.0 / d; }

// Wave-wide maximum of an unsigned value (DPP reduction of the device library).
extern "C" __device__ unsigned __ockl_wfred_max_u32(unsigned);

// Order-preserving 32-bit key of a non-negative magnitude; 0 is reserved for
// "no candidate".  Doubles keep their sign, exponent and 20 leading mantissa
// bits: rows closer than 2^-20 in magnitude tie and the lower one wins.
__device__ __forceinline__ unsigned mag_key(float m) { return __float_as_uint(m) + 1u; }
__device__ __forceinline__ unsigned mag_key(double m) {
    return (unsigned)((unsigned long long)__double_as_longlong(m) >> 32) + 1u;
}

// Exchange area of the register-resident elimination for NW waves.  A step
// takes two barriers: the waves' best keys first, then only the winning
// row.  (Every wave writing its candidate row before one barrier costs NW
// times the LDS writes, which serialise on the CU's LDS and cost more than
// the second barrier.)  Writes of a step happen after the previous step's
// last barrier and its reads before the next step's first, so one buffer
// suffices.
template <typename T, int NW>
struct ElimLds {
    alignas(16) T row[2 * PB];  // the pivot row (columns > the step): real parts, then imaginary
    T rp[2];                    // 1 / the pivot element
    unsigned long long kb[NW];  // best key of each wave (high word) and ~its thread (low word)
    int oid[PB], osrc[PB];      // chosen rows in pivot order
    int last;
    int zmin;  // first zero pivot of the root (1-based)
};

// Partial pivoting over the block's rows, one row per thread in registers
// (rows t >= nrows take no part; nrows >= PB), with implicit row
// interchanges: rows stay with their thread and a chosen row is only marked.
// Step jc: the largest |x_jc|^2 among the rows not chosen yet (by mag_key;
// ties: lower thread) becomes pivot jc, and the remaining rows are eliminated
// (multiplier stored in x_jc).  An all-zero column chooses the lowest open
// row and eliminates nothing.
// Afterwards the thread of pivot r holds row r of L11 \ U11 and returns r
// (-1 for rows not chosen); oid/osrc list the chosen rows in pivot order.
template <typename T, int NW>
__device__ __forceinline__ int elim_rows(ElimLds<T, NW>& L, T (&xr)[PB], T (&xi)[PB], int id, int src, int nrows) {
    const int t = threadIdx.x, wave = t >> 6;
    const bool valid = t < nrows;
    int mine = -1;
#pragma unroll
    for (int jc = 0; jc < PB; jc++) {
        const bool open = valid && mine < 0;
        const unsigned key = open ? mag_key(mag2(xr[jc], xi[jc])) : 0u;
        const unsigned wk = __ockl_wfred_max_u32(key);
        const int p = wave * 64 + __ffsll((long long)__ballot(key == wk)) - 1;  // ties: lowest thread
        if (t == p) L.kb[wave] = ((unsigned long long)wk << 32) | (0xffffffffu - (unsigned)p);
        __syncthreads();
        // largest key, ties to the lowest thread: one maximum over the packed words
        unsigned long long kbest = L.kb[0];
#pragma unroll
        for (int w = 1; w < NW; w++) kbest = max(kbest, L.kb[w]);
        const unsigned best = (unsigned)(kbest >> 32);
        const int pp = (int)(0xffffffffu - (unsigned)kbest);
        if (t == pp) {
            mine = jc;
            // 16-byte stores from the aligned piece holding column jc + 1
            using V = typename Vec16<T>::type;
            constexpr int E = 16 / sizeof(T);
#pragma unroll
            for (int q = 0; q < PB / E; q++)
                if (q * E + E - 1 > jc) {
                    V a, d;
                    __builtin_memcpy(&a, &xr[q * E], 16);
                    __builtin_memcpy(&d, &xi[q * E], 16);
                    *reinterpret_cast<V*>(L.row + q * E) = a;
                    *reinterpret_cast<V*>(L.row + PB + q * E) = d;
                }
            // reciprocal of the pivot (the rows scale by it, as zgetf2 does)
            const T d = xr[jc] * xr[jc] + xi[jc] * xi[jc];
            const T inv = d > T(0) ? recip(d) : T(0);
            L.rp[0] = xr[jc] * inv;
            L.rp[1] = -xi[jc] * inv;
        }
        __syncthreads();
        // branch-free update: rows that are not eliminated use multiplier 0
        // (keeps one copy of the row in registers across the step)
        const bool upd = open && best > 1u && t != pp;
        const T qr = L.rp[0], qi = L.rp[1];
        const T lr = upd ? xr[jc] * qr - xi[jc] * qi : T(0);
        const T li = upd ? xr[jc] * qi + xi[jc] * qr : T(0);
        xr[jc] = upd ? lr : xr[jc];
        xi[jc] = upd ? li : xi[jc];
        // x_c -= l u_c as four fused multiply-adds per element, the pivot row
        // read in 16-byte pieces from the aligned piece holding column jc + 1
        const T nlr = -lr, nli = -li;
        constexpr int E = 16 / sizeof(T);
#pragma unroll
        for (int q = 0; q < PB / E; q++)
            if (q * E + E - 1 > jc) {
                T ur[E], ui[E];
                const typename Vec16<T>::type a = *reinterpret_cast<const typename Vec16<T>::type*>(L.row + q * E);
                const typename Vec16<T>::type d = *reinterpret_cast<const typename Vec16<T>::type*>(L.row + PB + q * E);
                __builtin_memcpy(ur, &a, 16);
                __builtin_memcpy(ui, &d, 16);
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int c = q * E + e;
                    if (c > jc) {
                        xr[c] = fmadd(nlr, ur[e], xr[c]);
                        xr[c] = fmadd(li, ui[e], xr[c]);
                        xi[c] = fmadd(nlr, ui[e], xi[c]);
                        xi[c] = fmadd(nli, ur[e], xi[c]);
                    }
                }
            }
    }
    if (mine >= 0) {
        L.oid[mine] = id;
        L.osrc[mine] = src;
    }
    __syncthreads();
    return mine;
}

// Root of the tournament, one wave: pivot order -> LAPACK-style swaps
// ipiv[j0 .. j0+31], and their net effect on the touched rows as a list of
// (destination, source) row pairs: pl[0] = count, pl[1 + 2k] = destination,
// pl[2 + 2k] = source.  Lane l tracks one touched row (<= 64 rows).
__device__ void root_swaps(const int* __restrict__ oid, int j0, int* __restrict__ ipiv, int* __restrict__ pl) {
    const int lane = threadIdx.x;  // wave 0
    // pos: lane c < PB holds the current location of the c-th chosen row
    int pos = lane < PB ? oid[lane] : -1;
    int rows = -1, cont = -1, nt = 0;
    for (int c = 0; c < PB; c++) {
        const int x = j0 + c;
        const int y = __builtin_amdgcn_readlane(pos, c);
        if (lane == 0) ipiv[x] = y;
        if (lane > c && pos == x) pos = y;  // the row at x moves to y
        if (y == x) continue;
        // net permutation of the touched rows
        unsigned long long m = __ballot(rows == x);
        int ix;
        if (m == 0ull) {
            ix = nt++;
            if (lane == ix) rows = cont = x;
        } else {
            ix = __ffsll((long long)m) - 1;
        }
        m = __ballot(rows == y);
        int iy;
        if (m == 0ull) {
            iy = nt++;
            if (lane == iy) rows = cont = y;
        } else {
            iy = __ffsll((long long)m) - 1;
        }
        const int cx = __builtin_amdgcn_readlane(cont, ix), cy = __builtin_amdgcn_readlane(cont, iy);
        if (lane == ix) cont = cy;
        if (lane == iy) cont = cx;
    }
    const bool moved = lane < nt && cont != rows;
    const unsigned long long mm = __ballot(moved);
    const int k = __popcll(mm & ((1ull << lane) - 1ull));
    if (moved) {
        pl[1 + 2 * k] = rows;
        pl[2 + 2 * k] = cont;
    }
    if (lane == 0) pl[0] = __popcll(mm);
}

// ---------------------------------------------------------------------------
// Pivot tournament for columns j0..j0+31 over rows j0..n-1, one launch.
// Candidate sets of level L live at set index off[L] + i; a set is 32 ids and
// the 32 x 32 original values per plane.  cnt: one arrival counter per group
// per level (zero on entry; reset by the merging block).  The root writes
// ipiv, the row-pair list pl, the factored diagonal block d11 (32 x 32 per
// plane) and the first zero pivot into *info.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(Tslu<T>::CH) void k_tslu(const T* __restrict__ Ar, const T* __restrict__ Ai, long ld,
                                                      int j0, int n, int* __restrict__ cid, T* __restrict__ cvr,
                                                      T* __restrict__ cvi, int* __restrict__ cnt,
                                                      int* __restrict__ ipiv, int* __restrict__ pl,
                                                      T* __restrict__ d11r, T* __restrict__ d11i,
                                                      int* __restrict__ info) {
    __builtin_amdgcn_s_setprio(3);  // latency-bound: win issue slots over co-resident GEMM waves
    constexpr int CH = Tslu<T>::CH, G = CH / PB;
    __shared__ ElimLds<T, CH / 64> L;
    const int t = threadIdx.x;
    // level 0: rows j0 + b*CH .. (n - j0 is a multiple of PB, so every set is full)
    const int r0 = j0 + blockIdx.x * CH;
    int nrows = min(CH, n - r0);
    T xr[PB], xi[PB];
    if (t < nrows) {
        load_row(Ar + (long)(r0 + t) * ld + j0, xr);
        load_row(Ai + (long)(r0 + t) * ld + j0, xi);
    }
    int mine = elim_rows(L, xr, xi, r0 + t, r0 + t, nrows);
    int level_sets = (n - j0 + CH - 1) / CH;  // sets at this level
    int set = blockIdx.x, off = 0, cnt_off = 0;
    bool from_matrix = true;
    while (level_sets > 1) {
        // candidate set = original values of the chosen rows
        for (int e = t; e < PB * PB; e += CH) {
            const int r = e / PB, c = e % PB;
            const long o = ((long)(off + set) * PB + r) * PB + c;
            if (from_matrix) {
                const long g = (long)L.osrc[r] * ld + j0 + c;
                cvr[o] = Ar[g];
                cvi[o] = Ai[g];
            } else {
                const long g = (long)L.osrc[r] * PB + c;
                cvr[o] = cvr[g];
                cvi[o] = cvi[g];
            }
        }
        if (t < PB) cid[(off + set) * PB + t] = L.oid[t];
        // arrive at the group counter; the last arrival merges the group
        const int grp = set / G;
        const int gsize = min(G, level_sets - grp * G);
        __threadfence();
        __syncthreads();
        if (t == 0) {
            const int prev = atomicAdd(&cnt[cnt_off + grp], 1);
            L.last = prev == gsize - 1;
            if (L.last) cnt[cnt_off + grp] = 0;  // ready for the next launch
        }
        __syncthreads();
        if (!L.last) return;
        __threadfence();
        const int s0 = off + grp * G;
        nrows = gsize * PB;
        int id = -1;
        if (t < nrows) {
            load_row(cvr + ((long)s0 * PB + t) * PB, xr);
            load_row(cvi + ((long)s0 * PB + t) * PB, xi);
            id = cid[s0 * PB + t];
        }
        mine = elim_rows(L, xr, xi, id, s0 * PB + t, nrows);
        from_matrix = false;
        cnt_off += (level_sets + G - 1) / G;
        off += level_sets;
        set = grp;
        level_sets = (level_sets + G - 1) / G;
    }
    // root: the factored diagonal block, swaps, row-pair list, zero pivot
    if (t == 0) L.zmin = 0x7fffffff;
    __syncthreads();
    if (mine >= 0) {
        store_row(d11r + mine * PB, xr);
        store_row(d11i + mine * PB, xi);
        T dr = T(0), di = T(0);
#pragma unroll
        for (int c = 0; c < PB; c++)
            if (c == mine) {
                dr = xr[c];
                di = xi[c];
            }
        if (dr == T(0) && di == T(0)) atomicMin(&L.zmin, j0 + mine + 1);
    }
    __syncthreads();
    if (t == 0 && L.zmin != 0x7fffffff && *info == 0) *info = L.zmin;
    if (t < 64) root_swaps(L.oid, j0, ipiv, pl);
}

// Applies nl row-pair lists in order (list l at pl + l * PLS, each the net
// effect of one panel step's swaps, at most 2 PB pairs) to the columns
// [ca, cb) and [cc, cd).  A block owns 32 columns and 8 row groups; each
// thread reads the sources of its pairs of a list into registers (all loads
// in flight together), the block synchronises, and the thread writes them to
// the destinations, so every source is read before any destination of the
// list is written.  No LDS, so the kernel fits beside a running GEMM.
constexpr int PLS = 256;  // ints per row-pair list
constexpr int LASWP_COLS = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* __restrict__ Ar, T* __restrict__ Ai, long ld, int ca, int cb, int cc,
                                               int cd, const int* __restrict__ pl, int nl) {
    __builtin_amdgcn_s_setprio(3);
    constexpr int RG = 256 / LASWP_COLS;  // row groups
    constexpr int PER = 2 * PB / RG;      // pairs per thread and list
    const int tx = threadIdx.x % LASWP_COLS, ty = threadIdx.x / LASWP_COLS;
    const int v = blockIdx.x * LASWP_COLS + tx;  // virtual column over both ranges
    const int w1 = cb - ca;
    const int c = v < w1 ? ca + v : cc + (v - w1);
    const bool ok = v < w1 + (cd - cc);
    for (int l = 0; l < nl; l++) {
        const int* q = pl + l * PLS;
        const int np = q[0];
        if (np == 0) continue;
        T vr[PER], vi[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int k = ty + RG * u;
            if (ok && k < np) {
                const long o = (long)q[2 + 2 * k] * ld + c;
                vr[u] = Ar[o];
                vi[u] = Ai[o];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int k = ty + RG * u;
            if (ok && k < np) {
                const long o = (long)q[1 + 2 * k] * ld + c;
                Ar[o] = vr[u];
                Ai[o] = vi[u];
            }
        }
        __syncthreads();
    }
}

// Panel step after the swaps, with the factored 32 x 32 diagonal block d11
// from the tournament.  Block 0 writes d11 in place and forms U12 = L11^{-1}
// A12 for the nright columns right of the block (one column per thread);
// blocks >= 1 form L21 = A21 U11^{-1} for 256 rows each (one row per thread).
// Both substitutions run right-looking with the vector in registers and the
// triangle streamed from LDS in 16-byte chunks of one row (U) or one column
// (L, stored transposed), which keeps the kernel near 100 registers so it can
// share CUs with the trailing-update GEMM.  A zero pivot leaves its column
// unscaled.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(T* __restrict__ Ar, T* __restrict__ Ai, long ld, int j0, int nrows,
                                               int nright, const T* __restrict__ d11r,
                                               const T* __restrict__ d11i) {
    __builtin_amdgcn_s_setprio(3);
    using V = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
    __shared__ __align__(16) T ur[PB][PB], ui[PB][PB];  // U11 (rows)
    __shared__ __align__(16) T lr[PB][PB], li[PB][PB];  // L11 transposed: lr[l][r] = L[r][l]
    __shared__ T rr[PB], ri[PB];
    const int t = threadIdx.x;
    for (int e = t; e < PB * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        ur[r][c] = d11r[e];
        ui[r][c] = d11i[e];
        lr[c][r] = d11r[e];
        li[c][r] = d11i[e];
    }
    if (t < PB) {
        const T a = d11r[t * PB + t], b = d11i[t * PB + t];
        if (a == T(0) && b == T(0)) {
            rr[t] = T(1);
            ri[t] = T(0);
        } else {
            cdiv(T(1), T(0), a, b, rr[t], ri[t]);
        }
    }
    __syncthreads();
    // x_c -= s * M[l][c] for c > l (four fused multiply-adds per element),
    // row l of M streamed from LDS
    auto eliminate = [&](T (&xr)[PB], T (&xi)[PB], const T (*mr)[PB], const T (*mi)[PB], int l, T sr, T si) {
        const V* vr = reinterpret_cast<const V*>(&mr[l][0]);
        const V* vi = reinterpret_cast<const V*>(&mi[l][0]);
        const T nsr = -sr, nsi = -si;
#pragma unroll
        for (int q = 0; q < PB / E; q++) {
            if (q * E + E - 1 <= l) continue;
            T a[E], b[E];
            const V x = vr[q], y = vi[q];
            __builtin_memcpy(a, &x, 16);
            __builtin_memcpy(b, &y, 16);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int c = q * E + e;
                if (c > l) {
                    xr[c] = fmadd(nsr, a[e], xr[c]);
                    xr[c] = fmadd(si, b[e], xr[c]);
                    xi[c] = fmadd(nsr, b[e], xi[c]);
                    xi[c] = fmadd(nsi, a[e], xi[c]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    if (blockIdx.x == 0) {
        for (int e = t; e < PB * PB; e += 256) {
            const int r = e / PB, c = e % PB;
            Ar[(long)(j0 + r) * ld + j0 + c] = ur[r][c];
            Ai[(long)(j0 + r) * ld + j0 + c] = ui[r][c];
        }
        // Row r of the strip at a block-uniform address, the column as a
        // 32-bit offset; the store addresses are formed anew after the
        // substitution (the empty asm hides the offset from common-
        // subexpression elimination), else the 64 per-thread 64-bit load
        // addresses stay live across it (222 instead of ~100 registers).
        T* __restrict__ sr0 = Ar + (long)j0 * ld + j0 + PB;
        T* __restrict__ si0 = Ai + (long)j0 * ld + j0 + PB;
        for (int c = t; c < nright; c += 256) {
            T xr[PB], xi[PB];
#pragma unroll
            for (int r = 0; r < PB; r++) {
                xr[r] = (sr0 + r * ld)[(unsigned)c];
                xi[r] = (si0 + r * ld)[(unsigned)c];
            }
#pragma clang loop unroll(full)
            for (int l = 0; l < PB - 1; l++) eliminate(xr, xi, lr, li, l, xr[l], xi[l]);  // unit lower
            unsigned cs = (unsigned)c;
            asm volatile("" : "+v"(cs));
#pragma unroll
            for (int r = 0; r < PB; r++) {
                (sr0 + r * ld)[cs] = xr[r];
                (si0 + r * ld)[cs] = xi[r];
            }
        }
        return;
    }
    const int row = j0 + PB + (blockIdx.x - 1) * 256 + t;
    if (row >= j0 + nrows) return;
    T xr[PB], xi[PB];
    T* pr = Ar + (long)row * ld + j0;
    T* pi = Ai + (long)row * ld + j0;
    load_row(pr, xr);
    load_row(pi, xi);
#pragma clang loop unroll(full)  // a partial unroll indexes the row registers at run time
    for (int l = 0; l < PB; l++) {
        const T a = xr[l], b = xi[l];
        xr[l] = a * rr[l] - b * ri[l];  // x_l / U[l][l]
        xi[l] = a * ri[l] + b * rr[l];
        if (l < PB - 1) eliminate(xr, xi, ur, ui, l, xr[l], xi[l]);
    }
    store_row(pr, xr);
    store_row(pi, xi);
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

// U12 (NB x nc in the work buffer, ld lds) into the matrix rows (ld ldd) and
// its transpose Ut (nc x NB, ld NB) for the FP32 trailing update, whose
// operands are both k-contiguous (gemm_nt.h).  32 x 32 tiles through LDS;
// grid (nc / 32, NB / 32).
__global__ __launch_bounds__(256) void k_u12_out(const float* __restrict__ Sr, const float* __restrict__ Si, long lds,
                                                 float* __restrict__ Dr, float* __restrict__ Di, long ldd,
                                                 float* __restrict__ Utr, float* __restrict__ Uti) {
    __shared__ float tr[32][33], ti[32][33];
    const int c0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int r = ty; r < 32; r += 8) {
        const float vr = Sr[(long)(k0 + r) * lds + c0 + tx], vi = Si[(long)(k0 + r) * lds + c0 + tx];
        Dr[(long)(k0 + r) * ldd + c0 + tx] = vr;
        Di[(long)(k0 + r) * ldd + c0 + tx] = vi;
        tr[r][tx] = vr;
        ti[r][tx] = vi;
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < 32; r += 8) {
        Utr[(long)(c0 + r) * NB + k0 + tx] = tr[tx][r];
        Uti[(long)(c0 + r) * NB + k0 + tx] = ti[tx][r];
    }
}

// Work buffers of one factorisation (device).
template <typename T>
struct LuWork {
    int* info;       // first zero pivot (1-based), 0 if none
    int* cnt;        // tournament arrival counters
    int* cid;        // candidate row ids
    T *cvr, *cvi;    // candidate values
    int* pl;         // row-pair lists of the 8 steps of an outer panel (PLS ints each), two sets
    T *d11r, *d11i;  // factored 32 x 32 diagonal block
    T *ubr, *ubi;    // 256 x n U12 buffer
    T *utr, *uti;    // n x 256 transpose of U12 (FP32 trailing update)
    T *tr, *ti;      // 128 x 128 product of the L11^{-1} doubling
};

// The row-pair lists of outer panel k0.  Two sets, by panel parity: the
// update stream applies panel K's swaps to the far columns while the
// tournaments of panel K+1 already write theirs.
template <typename T>
int* panel_lists(const LuWork<T>& w, int k0) {
    return w.pl + ((k0 / NB) & 1) * (NB / PB) * PLS;
}

template <typename T>
void launch_laswp(bem_ctx* c, hipStream_t s, T* Ar, T* Ai, long ld, int ca, int cb, int cc, int cd, const int* pl,
                  int nl) {
    const int ncols = (cb - ca) + (cd - cc);
    if (ncols <= 0) return;
    const int kt = kt_mark(c, KS_LASWP, s);
    hipLaunchKernelGGL(k_laswp<T>, dim3((ncols + LASWP_COLS - 1) / LASWP_COLS), dim3(256), 0, s, Ar, Ai, ld, ca, cb,
                       cc, cd, pl, nl);
    kt_close(c, kt, 0.0, 2.0 * 2.0 * sizeof(T) * (double)ncols * 2 * PB * nl);
}

// L11^{-1} of the NB x NB unit lower triangle at F (ld ldf) into I (ld NB) by
// doubling.  The inverses of the eight 32 x 32 diagonal blocks come from one
// launch (one wave per block, k_inv_lower32); then for s = 32, 64, 128 every pair of s x s
// diagonal blocks is joined,
//     [X11 0; X21 X22] with X21 = -X22 (L21 X11),
// as two batched GEMMs (all pairs of a level in one launch each, on the 32 x 32
// tile: many blocks, little LDS, since they run beside the trailing update).  The blocks
// above the diagonal are not written: the caller keeps them zero.  Seven
// launches with short dependency chains replace one 256-step substitution.
template <typename T>
void invert_lower_doubling(hipStream_t s, const T* Fr, const T* Fi, long ldf, T* Ir, T* Ii, const LuWork<T>& w) {
    constexpr int D = 32;
    hipLaunchKernelGGL(k_inv_lower32<T>, dim3(NB / D), dim3(64), 0, s, Fr, Fi, ldf, (long)D * (ldf + 1), Ir, Ii,
                       (long)NB, (long)D * (NB + 1));
    for (int h = D; h < NB; h *= 2) {
        const int pairs = NB / (2 * h);
        const long sf = 2L * h * (ldf + 1), si = 2L * h * (NB + 1), st = (long)h * h;
        const long f21 = (long)h * ldf, i21 = (long)h * NB, i22 = (long)h * (NB + 1);
        // T = L21 X11, then X21 = -X22 T
        if constexpr (sizeof(T) == 4) {  // register-direct short-K products (no LDS, no barrier)
            if (cgemm_rd_fits(h, h, h, Fr + f21, Fi + f21, ldf, w.tr, w.ti, h, sf, st) &&
                cgemm_rd_fits(h, h, h, Ir + i22, Ii + i22, NB, Ir + i21, Ii + i21, NB, si, si)) {
                cgemm_rd(s, h, h, h, Fr + f21, Fi + f21, ldf, Ir, Ii, NB, w.tr, w.ti, h, 1 | GEMM_MODE_URGENT, pairs,
                         sf, si, st);
                cgemm_rd(s, h, h, h, Ir + i22, Ii + i22, NB, w.tr, w.ti, h, Ir + i21, Ii + i21, NB,
                         3 | GEMM_MODE_URGENT, pairs, si, st, si);
                continue;
            }
        }
        cgemm_any<T>(kGemmSmall, s, h, h, h, Fr + f21, Fi + f21, ldf, Ir, Ii, NB, w.tr, w.ti, h, 1 | GEMM_MODE_URGENT,
                     pairs, sf, si, st);
        cgemm_any<T>(kGemmSmall, s, h, h, h, Ir + i22, Ii + i22, NB, w.tr, w.ti, h, Ir + i21, Ii + i21, NB,
                     3 | GEMM_MODE_URGENT, pairs, si, st,
                     si);
    }
}

// Factors the outer panel of columns k0 .. k0+NB-1 (rows k0 .. n-1) on stream
// s: 8 steps of tournament, swaps within the panel's columns, L21/U12 of the
// step and the rank-32 update of the panel's remaining columns.  The swaps of
// step j are kept as row-pair list j for the columns outside the panel.
// Look-ahead inside the panel: the rank-32 update of the next 32 columns
// stays on s (the next tournament needs them); the update of the columns
// beyond runs on stream ss and is joined before the next step's swaps, so
// every column still receives its updates in step order.
template <typename T>
void factor_panel(bem_ctx* c, hipStream_t s, hipStream_t ss, T* Ar, T* Ai, int k0, const LuWork<T>& w) {
    const int n = c->npad;
    const long ld = n;
    hipEvent_t e_step = c->lu_ev[4], e_rest = c->lu_ev[5];
    bool rest_pending = false;
    for (int j0 = k0; j0 < k0 + NB; j0 += PB) {
        const int m = n - j0;
        const int nsets = (m + Tslu<T>::CH - 1) / Tslu<T>::CH;
        int* pl = panel_lists(w, k0) + ((j0 - k0) / PB) * PLS;
        int kt = kt_mark(c, KS_TSLU, s);
        hipLaunchKernelGGL(k_tslu<T>, dim3(nsets), dim3(Tslu<T>::CH), 0, s, Ar, Ai, ld, j0, n, w.cid, w.cvr, w.cvi, w.cnt,
                           c->ipiv, pl, w.d11r, w.d11i, w.info);
        kt_close(c, kt, 0.0, 2.0 * sizeof(T) * (double)m * PB);
        if (rest_pending) {  // the previous step's update of the far columns
            (void)hipStreamWaitEvent(s, e_rest, 0);
            rest_pending = false;
        }
        launch_laswp<T>(c, s, Ar, Ai, ld, k0, k0 + NB, 0, 0, pl, 1);
        const int nright = k0 + NB - (j0 + PB);
        const int nblk = 1 + (m - PB + 255) / 256;
        kt = kt_mark(c, KS_PANEL, s);
        hipLaunchKernelGGL(k_panel<T>, dim3(nblk), dim3(256), 0, s, Ar, Ai, ld, j0, m, nright, w.d11r, w.d11i);
        kt_close(c, kt, 4.0 * (double)m * PB * PB, 4.0 * sizeof(T) * (double)m * PB);
        if (nright > 0) {
            const long o21 = (long)(j0 + PB) * ld + j0, o12 = (long)j0 * ld + j0 + PB;
            const long o22 = (long)(j0 + PB) * ld + j0 + PB;
            if (nright > PB) {
                (void)hipEventRecord(e_step, s);  // L21 and U12 of this step are in place
                (void)hipStreamWaitEvent(ss, e_step, 0);
                cgemm_ctx<T>(c, m - PB, nright - PB, PB, Ar + o21, Ai + o21, ld, Ar + o12 + PB, Ai + o12 + PB, ld,
                             Ar + o22 + PB, Ai + o22 + PB, ld, 0, ss, GEMM_PANEL);
                (void)hipEventRecord(e_rest, ss);
                rest_pending = true;
            }
            cgemm_ctx<T>(c, m - PB, PB, PB, Ar + o21, Ai + o21, ld, Ar + o12, Ai + o12, ld, Ar + o22, Ai + o22, ld, 0,
                         s, GEMM_PANEL);
        }
    }
}

// U12 = L11^{-1} A12 and A22 -= L21 U12 for the columns [c0, c0 + nc) right
// of outer panel k0, on stream us; U12 goes through the work buffer.  FP32:
// U12 as a GEMM with the explicit L11^{-1} (Lir, Lii); FP64: by substitution
// with L11 itself (k_trsm_rhs), the backward-stable form of Zgetrf's TRSM.
template <typename T>
void update_columns(bem_ctx* c, hipStream_t us, T* Ar, T* Ai, const T* Lir, const T* Lii, int k0, int c0, int nc,
                    const LuWork<T>& w, bool headroom) {
    const int n = c->npad;
    const long ld = n;
    const int nr = n - (k0 + NB);
    const long io = (long)(k0 / NB) * NB * NB;
    const long o12 = (long)k0 * ld + c0, o21 = (long)(k0 + NB) * ld + k0, o22 = (long)(k0 + NB) * ld + c0;
    T* ubr = w.ubr + (c0 - (k0 + NB));
    T* ubi = w.ubi + (c0 - (k0 + NB));
    if constexpr (sizeof(T) == 8) {
        const long o11 = (long)k0 * ld + k0;
        const int kt = kt_mark(c, KS_TRSM, us);
        trsm_rhs<T>(us, true, Ar + o11, Ai + o11, ld, Ar + o12, Ai + o12, ld, ubr, ubi, ld, nc);
        kt_close(c, kt, 4.0 * (double)NB * NB * nc, 2.0 * sizeof(T) * ((double)NB * NB + 2.0 * NB * nc));
    } else {
        cgemm_ctx<T>(c, NB, nc, NB, Lir + io, Lii + io, NB, Ar + o12, Ai + o12, ld, ubr, ubi, ld, 1, us);
        // FP32: the trailing update on the k-contiguous kernel (U12 transposed
        // beside its copy into the matrix), when the shape fits it
        const GemmPlan& plan = c->gemm_plan[0];
        float* utr = w.utr + (long)(c0 - (k0 + NB)) * NB;
        float* uti = w.uti + (long)(c0 - (k0 + NB)) * NB;
        if (cupdate_nt_fits(plan.nt, nr, nc, NB, Ar + o21, Ai + o21, ld, utr, uti, NB, Ar + o22, Ai + o22, ld)) {
            int kt = kt_mark(c, KS_OTHER, us);
            hipLaunchKernelGGL(k_u12_out, dim3((unsigned)(nc / 32), NB / 32), dim3(256), 0, us, ubr, ubi, ld,
                               Ar + o12, Ai + o12, ld, utr, uti);
            kt_close(c, kt, 0.0, 6.0 * sizeof(T) * (double)NB * nc);
            kt = kt_mark(c, KS_UPDATE32, us);
            cupdate_nt(plan.nt, us, nr, nc, NB, Ar + o21, Ai + o21, ld, utr, uti, NB, Ar + o22, Ai + o22, ld,
                       headroom ? gemm_headroom_cap(plan, plan.large) : 0);
            kt_close(c, kt, 6.0 * (double)nr * nc * NB,
                     2.0 * sizeof(T) * ((double)nr * NB + (double)NB * nc + 2.0 * nr * nc));
            return;
        }
    }
    cgemm_ctx<T>(c, nr, nc, NB, Ar + o21, Ai + o21, ld, ubr, ubi, ld, Ar + o22, Ai + o22, ld, 0, us,
                 GEMM_UPDATE | (headroom ? GEMM_HEADROOM : 0));
    const int kt = kt_mark(c, KS_OTHER, us);
    hipLaunchKernelGGL(k_copy_rows<T>, dim3((unsigned)(((long)NB * nc + 255) / 256)), dim3(256), 0, us, NB, nc, ubr,
                       ubi, ld, Ar + o12, Ai + o12, ld);
    kt_close(c, kt, 0.0, 4.0 * sizeof(T) * (double)NB * nc);
}

// Blocked right-looking factorisation of the padded npad x npad matrix with
// look-ahead of depth one: outer panel K+1 is factored on the panel stream
// while the trailing update of panel K runs on the update stream.  Panel K+1
// only touches its own columns; its swaps reach the other columns on the
// update stream.  After panel K, the critical path to panel K+1 is: swaps in
// the columns of panel K+1 (update stream) beside L11^{-1} (panel stream),
// then U12 and the update of those 256 columns; the swaps and the update of
// all other columns follow on the update stream while panel K+1 is factored.
// Stream layout (c->lu_streams): 0 = everything on the context stream (no
// overlap); 1 = panel stream and context stream over all CUs.
template <typename T>
void factor(bem_ctx* c, T* Ar, T* Ai, T* Lir, T* Lii, T* Uir, T* Uii, const LuWork<T>& w) {
    const int n = c->npad, nb = n / NB;
    const long ld = n;
    hipStream_t us = c->stream, ps = c->stream2, ss = c->stream3;
    if (c->lu_streams == 0) ps = ss = us;
    hipEvent_t e_panel = c->lu_ev[0], e_next = c->lu_ev[1], e_start = c->lu_ev[2], e_linv = c->lu_ev[3];
    (void)hipEventRecord(e_start, c->stream);  // the matrix and work buffers are ready
    (void)hipStreamWaitEvent(ps, e_start, 0);
    if (us != c->stream) (void)hipStreamWaitEvent(us, e_start, 0);
    factor_panel<T>(c, ps, ss, Ar, Ai, 0, w);
    for (int K = 0; K < nb; K++) {
        const int k0 = K * NB;
        const int nr = n - (k0 + NB);
        const long o11 = (long)k0 * ld + k0;
        const long io = (long)K * NB * NB;
        (void)hipEventRecord(e_panel, ps);
        if constexpr (sizeof(T) == 4) {  // FP64 solves with L11 by substitution instead
            const int kt = kt_mark(c, KS_TRSM, ps);
            invert_lower_doubling<T>(ps, Ar + o11, Ai + o11, ld, Lir + io, Lii + io, w);
            kt_close(c, kt, 8.0 * NB * NB * NB / 6.0, 4.0 * sizeof(T) * (double)NB * NB);
        }
        (void)hipEventRecord(e_linv, ps);
        (void)hipStreamWaitEvent(us, e_panel, 0);
        if (nr <= 0) {
            launch_laswp<T>(c, us, Ar, Ai, ld, 0, k0, 0, 0, panel_lists(w, k0), NB / PB);
            (void)hipStreamWaitEvent(us, e_linv, 0);
            break;
        }
        // columns of panel K+1 first, then hand panel K+1 to the panel stream
        launch_laswp<T>(c, us, Ar, Ai, ld, k0 + NB, k0 + 2 * NB, 0, 0, panel_lists(w, k0), NB / PB);
        (void)hipStreamWaitEvent(us, e_linv, 0);
        update_columns<T>(c, us, Ar, Ai, Lir, Lii, k0, k0 + NB, NB, w, false);
        (void)hipEventRecord(e_next, us);
        (void)hipStreamWaitEvent(ps, e_next, 0);
        // the rest of the update is enqueued before panel K+1's ~40 launches,
        // so a host that enqueues slowly (under a profiler, say) does not
        // hold the update stream back
        launch_laswp<T>(c, us, Ar, Ai, ld, 0, k0, k0 + 2 * NB, n, panel_lists(w, k0), NB / PB);
        if (nr > NB)
            update_columns<T>(c, us, Ar, Ai, Lir, Lii, k0, k0 + 2 * NB, nr - NB, w, ps != us && c->gemm_plan[sizeof(T) == 4 ? 0 : 1].cap > 0);
        factor_panel<T>(c, ps, ss, Ar, Ai, k0 + NB, w);
    }
    // U^{-1} of all diagonal blocks in one launch (FP32; the FP64 solves
    // substitute with the diagonal blocks themselves)
    if constexpr (sizeof(T) == 4) {
        const int kt = kt_mark(c, KS_TRSM, us);
        invert_diag_blocks<T>(us, NB, nb, Ar, Ai, ld, Uir, Uii, false);
        kt_close(c, kt, 8.0 * (double)n * NB * NB / 6.0, 4.0 * sizeof(T) * (double)n * NB);
    }
    if (us != c->stream) {  // the context stream continues after the last update
        (void)hipEventRecord(e_start, us);
        (void)hipStreamWaitEvent(c->stream, e_start, 0);
    }
}

template <typename T>
__global__ void k_convert(const double* __restrict__ a, T* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

template <typename T>
int factor_prec(bem_ctx* c, T* Ar, T* Ai, void*& inv, size_t& inv_bytes) {
    const int n = c->npad;
    // diagonal-block inverses (FP32 only): L and U, npad x NB per plane each
    constexpr bool with_inv = sizeof(T) == 4;
    const size_t ib = 4 * (size_t)n * NB * sizeof(T);
    if (with_inv && inv_bytes < ib) {
        if (inv) (void)hipFree(inv);
        inv = nullptr;
        inv_bytes = 0;
        if (hipMalloc(&inv, ib) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "LU: diagonal-block inverses");
        inv_bytes = ib;
    }
    T* Lir = with_inv ? static_cast<T*>(inv) : nullptr;
    T* Lii = with_inv ? Lir + (size_t)n * NB : nullptr;
    T* Uir = with_inv ? Lii + (size_t)n * NB : nullptr;
    T* Uii = with_inv ? Uir + (size_t)n * NB : nullptr;
    // work: info, counters, row-pair list, candidate ids, candidate values,
    // diagonal block, U12 buffer
    const int nsets0 = (n + Tslu<T>::CH - 1) / Tslu<T>::CH;
    const size_t ncnt = 2 * (size_t)nsets0 + 64;
    const size_t nint = 64 + ncnt + 2 * (NB / PB) * PLS + 2 * (size_t)nsets0 * PB;
    const size_t bytes = nint * sizeof(int) + 256 + 4 * (size_t)nsets0 * PB * PB * sizeof(T) +
                         2 * (size_t)PB * PB * sizeof(T) + 4 * (size_t)NB * n * sizeof(T) +
                         2 * (size_t)(NB / 2) * (NB / 2) * sizeof(T) + 256;
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    LuWork<T> w;
    w.info = static_cast<int*>(scr);
    w.cnt = w.info + 64;
    w.pl = w.cnt + ncnt;
    w.cid = w.pl + 2 * (NB / PB) * PLS;
    char* p = reinterpret_cast<char*>(w.info + nint);
    p += (256 - reinterpret_cast<uintptr_t>(p) % 256) % 256;
    w.cvr = reinterpret_cast<T*>(p);
    w.cvi = w.cvr + 2 * (size_t)nsets0 * PB * PB;
    w.d11r = w.cvi + 2 * (size_t)nsets0 * PB * PB;
    w.d11i = w.d11r + PB * PB;
    w.ubr = w.d11i + PB * PB;
    w.ubi = w.ubr + (size_t)NB * n;
    w.utr = w.ubi + (size_t)NB * n;
    w.uti = w.utr + (size_t)NB * n;
    w.tr = w.uti + (size_t)NB * n;
    w.ti = w.tr + (size_t)(NB / 2) * (NB / 2);
    BEM_HIP(c, hipMemsetAsync(w.info, 0, (64 + ncnt) * sizeof(int), c->stream));
    // the doubling writes only the diagonal and lower blocks of each L11^{-1}
    if (with_inv) BEM_HIP(c, hipMemsetAsync(Lir, 0, 2 * (size_t)n * NB * sizeof(T), c->stream));
    factor<T>(c, Ar, Ai, Lir, Lii, Uir, Uii, w);
    return 0;
}

}  // namespace

int lu_setup_streams(bem_ctx* c) {
    if (c->lu_streams >= 0) return 0;
    // (A layout with the panel stream on CUs reserved by a CU mask ended in a
    // host-side crash in its first run, cause not established; it was removed.)
    const char* e1 = std::getenv("BEM_LU_STREAMS");
    const int mode = e1 ? std::atoi(e1) : 1;
    c->lu_streams = mode == 0 ? 0 : 1;
    return 0;
}

namespace {

// Systems from this size on choose the trailing-update tile by whole trial
// factorisations: how well the update GEMM overlaps the panel stream decides
// the factorisation time, and a GEMM timed alone does not show that.
constexpr int kTrialMinN = 8192;

// The working copy of precision T from the FP64 matrix Z.
template <typename T>
int load_working_copy(bem_ctx* c, T* Ar, T* Ai) {
    const size_t nn = (size_t)c->npad * c->npad;
    if constexpr (sizeof(T) == 8) {
        BEM_HIP(c, hipMemcpyAsync(Ar, c->Zr, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        BEM_HIP(c, hipMemcpyAsync(Ai, c->Zi, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else {
        const size_t blocks = (nn + 255) / 256;
        hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, c->stream, c->Zr, Ar, nn);
        hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, c->stream, c->Zi, Ai, nn);
    }
    return 0;
}

// One timed factorisation of the context's matrix per tile configuration;
// the fastest becomes the plan's trailing-update tile.  Every configuration
// gives the same factors (same accumulation order), so the trials change
// nothing but the choice.  Sets *ran when the working copy was overwritten.
template <typename T>
int factor_trials(bem_ctx* c, T* Ar, T* Ai, void*& inv, size_t& inv_bytes, bool* ran) {
    GemmPlan& plan = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    *ran = false;
    if (plan.trialed) return 0;
    if (c->npad < kTrialMinN) {
        plan.trialed = true;
        return 0;
    }
    if (!c->has_Z64) return 0;  // no source to restore the working copy from
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return bem_fail(c, BEM_ERR_HIP, "LU trials: events");
    }
    *ran = true;
    int rc = 0;
    // FP32 tries the configurations of its trailing-update kernel (k_cupdate_nt),
    // FP64 the large tile of k_cgemm
    constexpr bool f32 = sizeof(T) == 4;
    const int ncfg = f32 ? kNtCfgs : kGemmTrialCfgs;
    for (int cfg = 0; cfg < ncfg && !rc; cfg++) {
        (f32 ? plan.nt : plan.large) = cfg;
        if ((rc = load_working_copy<T>(c, Ar, Ai))) break;
        (void)hipEventRecord(e0, c->stream);
        if ((rc = factor_prec<T>(c, Ar, Ai, inv, inv_bytes))) break;
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = bem_fail(c, BEM_ERR_HIP, "LU trials: factorisation");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        plan.trial_ms[cfg] = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the fastest trial; among trials within 1 % of it (run-to-run noise),
    // FP64 takes the configuration with the higher rate timed alone, FP32 the
    // first (configuration 0 is the fastest alone).  (A 3 % band once took
    // the FP64 64 x 64 / K 8 tile at 368 ms over the 64 x 64 tile at 358 ms,
    // and its factorisations then took 372 ms against 361.)
    int best = 0;
    for (int cfg = 1; cfg < ncfg; cfg++)
        if (plan.trial_ms[cfg] < plan.trial_ms[best]) best = cfg;
    const double lim = 1.01 * plan.trial_ms[best];
    int pick = best;
    for (int cfg = 0; cfg < ncfg; cfg++) {
        if (plan.trial_ms[cfg] > lim) continue;
        if (f32 ? cfg < pick : plan.alone_tflops[cfg] > plan.alone_tflops[pick]) pick = cfg;
    }
    (f32 ? plan.nt : plan.large) = pick;
    if (rc) return rc;
    plan.trialed = true;
    gemm_cache_store(c, sizeof(T) == 4);
    return 0;
}

}  // namespace

int lu_factor_launch(bem_ctx* c, int precision) {
    const int n = c->npad;
    hipStream_t s = c->stream;
    int rc = gemm_ensure_tuned(c, precision != BEM_F64);
    if (rc) return rc;
    if ((rc = lu_setup_streams(c))) return rc;
    bool trials = false;
    if (precision == BEM_F64) {
        void* no_inv = nullptr;  // FP64 factors keep no diagonal-block inverses
        size_t no_inv_bytes = 0;
        if ((rc = factor_trials<double>(c, c->Vr, c->Vi, no_inv, no_inv_bytes, &trials))) return rc;
        if ((rc = load_working_copy<double>(c, c->Vr, c->Vi))) return rc;
        rc = factor_prec<double>(c, c->Vr, c->Vi, no_inv, no_inv_bytes);
    } else {
        if ((rc = factor_trials<float>(c, c->Wr, c->Wi, c->inv32, c->inv32_bytes, &trials))) return rc;
        if (!c->has_Z32 || trials)
            if ((rc = load_working_copy<float>(c, c->Wr, c->Wi))) return rc;
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
