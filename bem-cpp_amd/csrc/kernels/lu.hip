// lu.hip -- blocked complex LU with row pivoting on gfx950, FP64 and FP32.
//
// Replaces cuSOLVER Zgetrf (solver.cu:7-48, 87-113).  The reference factors
// the column-major view of its row-major Z (i.e. Z^T) and solves with OP_T;
// this library factors P Z = L U directly on the row-major matrix (rows are
// contiguous, so row interchanges are contiguous copies) and solves Z X = B.
// The pivots differ from cuSOLVER's; the solution does not.
//
// Right-looking two-level blocking on the padded matrix (npad = 256 k):
//   outer panel of NB = 256 columns, factored in 8 inner steps of B = 32:
//     1. pivot selection for the 32 columns over all remaining rows by a
//        tournament (communication-avoiding "TSLU": Grigori, Demmel & Xiang,
//        SIAM J. Matrix Anal. Appl. 32 (2011)): 256-row chunks each choose 32
//        candidate rows by partial pivoting in LDS, candidates are merged
//        8 sets at a time until one set remains; its pivot order becomes
//        LAPACK-style swaps ipiv[j0 .. j0+31]
//     2. the swaps are applied to whole rows (all columns)
//     3. the 32 x 32 diagonal block is factored without further pivoting,
//        L21 = A21 U11^{-1} row by row, U12 of the rest of the outer panel
//     4. rank-32 update of the rest of the outer panel
//   then U12 = L11^{-1} A12 (256 x n') and the rank-256 trailing update
//   A22 -= L21 U12 on the matrix cores (gemm.h; 3M form in FP32).
// Each step is a separate launch on the context stream: no grid-wide barrier
// and no atomics, so every wave always runs to completion.
#include "../bem_internal.h"
#include "gemm.h"
#include "trsm.h"

namespace {

constexpr int NB = 256;  // outer panel width
constexpr int PB = 32;   // inner panel width (pivot block)
constexpr int CH = 256;  // rows per tournament block
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

// ---------------------------------------------------------------------------
// Tournament step.  Loads up to CH candidate rows (32 columns each) either
// straight from the matrix (level 0: rows row0 + blockIdx.x*CH ..) or from
// the previous level's candidate sets (G sets per block), runs partial
// pivoting in LDS, and writes the chosen 32 rows (ids and original values) as
// candidate set blockIdx.x.  With `final`, also turns the pivot order into
// LAPACK-style swaps ipiv[j0 + c] (0-based row indices).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_tslu(int level, const T* __restrict__ Ar, const T* __restrict__ Ai,
                                              long ld, int j0, int row_end, int nsets_in,
                                              const int* __restrict__ cid_in, const T* __restrict__ cvr_in,
                                              const T* __restrict__ cvi_in, int* __restrict__ cid_out,
                                              T* __restrict__ cvr_out, T* __restrict__ cvi_out, int final_level,
                                              int* __restrict__ ipiv) {
    __shared__ T vr[CH][PB + 1], vi[CH][PB + 1];
    __shared__ int id[CH], src[CH];
    __shared__ T sv[4];
    __shared__ int si[4];
    const int t = threadIdx.x;
    int nrows;
    if (level == 0) {
        const int r0 = j0 + blockIdx.x * CH;
        nrows = min(CH, row_end - r0);
        if (t < nrows) {
            const long o = (long)(r0 + t) * ld + j0;
            for (int c = 0; c < PB; c++) {
                vr[t][c] = Ar[o + c];
                vi[t][c] = Ai[o + c];
            }
            id[t] = r0 + t;
            src[t] = r0 + t;
        }
    } else {
        const int s0 = blockIdx.x * G;
        const int ns = min(G, nsets_in - s0);
        nrows = ns * PB;
        if (t < nrows) {
            const long o = ((long)s0 * PB + t) * PB;
            for (int c = 0; c < PB; c++) {
                vr[t][c] = cvr_in[o + c];
                vi[t][c] = cvi_in[o + c];
            }
            id[t] = cid_in[s0 * PB + t];
            src[t] = s0 * PB + t;
        }
    }
    __syncthreads();
    const int nsel = min(PB, nrows);
    for (int jc = 0; jc < nsel; jc++) {
        const T v = (t >= jc && t < nrows) ? mag2(vr[t][jc], vi[t][jc]) : T(-1);
        T best;
        int p;
        argmax256(v, t, sv, si, best, p);
        if (p != jc && t < PB) {  // swap LDS rows jc and p
            T a = vr[jc][t]; vr[jc][t] = vr[p][t]; vr[p][t] = a;
            a = vi[jc][t]; vi[jc][t] = vi[p][t]; vi[p][t] = a;
        }
        if (p != jc && t == 0) {
            int x = id[jc]; id[jc] = id[p]; id[p] = x;
            x = src[jc]; src[jc] = src[p]; src[p] = x;
        }
        __syncthreads();
        if (best > T(0) && t > jc && t < nrows) {
            T lr, li;
            cdiv(vr[t][jc], vi[t][jc], vr[jc][jc], vi[jc][jc], lr, li);
            for (int c = jc + 1; c < PB; c++) {
                const T ur = vr[jc][c], ui = vi[jc][c];
                vr[t][c] -= lr * ur - li * ui;
                vi[t][c] -= lr * ui + li * ur;
            }
        }
        __syncthreads();
    }
    // chosen rows 0..nsel-1: ids and original values
    if (t < nsel) {
        cid_out[blockIdx.x * PB + t] = id[t];
        const long o = ((long)blockIdx.x * PB + t) * PB;
        if (level == 0) {
            const long g = (long)src[t] * ld + j0;
            for (int c = 0; c < PB; c++) {
                cvr_out[o + c] = Ar[g + c];
                cvi_out[o + c] = Ai[g + c];
            }
        } else {
            const long g = (long)src[t] * PB;
            for (int c = 0; c < PB; c++) {
                cvr_out[o + c] = cvr_in[g + c];
                cvi_out[o + c] = cvi_in[g + c];
            }
        }
    }
    if (final_level && t == 0) {
        // Pivot order -> sequential swaps: row j0+c is exchanged with the
        // current position of the c-th chosen row.  A chosen row that sits
        // at the position being vacated moves to the other side of the swap.
        int pos[PB];
        for (int c = 0; c < PB; c++) pos[c] = c < nsel ? id[c] : j0 + c;
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
// zero pivot in *info.  Blocks >= 1 compute L21 = A21 U11^{-1} for 256 rows each.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(T* __restrict__ Ar, T* __restrict__ Ai, long ld, int j0, int nrows,
                                               int nright, int* __restrict__ info) {
    __shared__ T ur[PB][PB + 1], ui[PB][PB + 1];
    const int t = threadIdx.x;
    for (int e = t; e < PB * PB; e += 256) {
        const int r = e / PB, c = e % PB;
        ur[r][c] = Ar[(long)(j0 + r) * ld + j0 + c];
        ui[r][c] = Ai[(long)(j0 + r) * ld + j0 + c];
    }
    __syncthreads();
    for (int k = 0; k < PB; k++) {
        const T pr = ur[k][k], pi = ui[k][k];
        const bool zero = pr == T(0) && pi == T(0);
        if (!zero && t > k && t < PB) {
            T lr, li;
            cdiv(ur[t][k], ui[t][k], pr, pi, lr, li);
            ur[t][k] = lr;
            ui[t][k] = li;
            for (int c = k + 1; c < PB; c++) {
                const T a = ur[k][c], b = ui[k][c];
                ur[t][c] -= lr * a - li * b;
                ui[t][c] -= lr * b + li * a;
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
        // U12 of the rest of the outer panel: unit-lower forward substitution per column
        for (int c = t; c < nright; c += 256) {
            const long base = (long)j0 * ld + j0 + PB + c;
            T xr[PB], xi[PB];
            for (int r = 0; r < PB; r++) {
                T sr = Ar[base + (long)r * ld], si = Ai[base + (long)r * ld];
                for (int l = 0; l < r; l++) {
                    sr -= ur[r][l] * xr[l] - ui[r][l] * xi[l];
                    si -= ur[r][l] * xi[l] + ui[r][l] * xr[l];
                }
                xr[r] = sr;
                xi[r] = si;
            }
            for (int r = 0; r < PB; r++) {
                Ar[base + (long)r * ld] = xr[r];
                Ai[base + (long)r * ld] = xi[r];
            }
        }
        return;
    }
    // L21 rows: x U11 = a  ->  x_c = (a_c - sum_{l<c} x_l U[l][c]) / U[c][c]
    const int row = j0 + PB + (blockIdx.x - 1) * 256 + t;
    if (row >= j0 + nrows) return;
    const long o = (long)row * ld + j0;
    T xr[PB], xi[PB];
    for (int c = 0; c < PB; c++) {
        T sr = Ar[o + c], si = Ai[o + c];
        for (int l = 0; l < c; l++) {
            sr -= xr[l] * ur[l][c] - xi[l] * ui[l][c];
            si -= xr[l] * ui[l][c] + xi[l] * ur[l][c];
        }
        T qr = sr, qi = si;
        if (!(ur[c][c] == T(0) && ui[c][c] == T(0))) cdiv(sr, si, ur[c][c], ui[c][c], qr, qi);
        xr[c] = qr;
        xi[c] = qi;
    }
    for (int c = 0; c < PB; c++) {
        Ar[o + c] = xr[c];
        Ai[o + c] = xi[c];
    }
}

}  // namespace

namespace {

// Blocked factorisation of the padded npad x npad matrix at (Ar, Ai).
template <typename T>
int factor(bem_ctx* c, T* Ar, T* Ai, int* dinfo, int* cid, T* cvr, T* cvi) {
    const int n = c->npad;
    const long ld = n;
    hipStream_t s = c->stream;
    for (int k0 = 0; k0 < n; k0 += NB) {
        for (int j0 = k0; j0 < k0 + NB; j0 += PB) {
            const int m = n - j0;
            // tournament over the rows j0..n-1
            int nsets = (m + CH - 1) / CH;
            int level = 0;
            int* ci = cid;
            T *vr = cvr, *vi = cvi;
            int* co = cid + (size_t)n / PB * PB;  // second half of the candidate buffers
            T *wr = cvr + (size_t)n * PB, *wi = cvi + (size_t)n * PB;
            int kt = kt_mark(c, KS_TSLU);
            hipLaunchKernelGGL(k_tslu<T>, dim3(nsets), dim3(256), 0, s, 0, Ar, Ai, ld, j0, n, 0, nullptr, nullptr,
                               nullptr, ci, vr, vi, nsets == 1 ? 1 : 0, c->ipiv);
            while (nsets > 1) {
                const int nout = (nsets + G - 1) / G;
                level++;
                hipLaunchKernelGGL(k_tslu<T>, dim3(nout), dim3(256), 0, s, level, Ar, Ai, ld, j0, n, nsets, ci, vr,
                                   vi, co, wr, wi, nout == 1 ? 1 : 0, c->ipiv);
                std::swap(ci, co);
                std::swap(vr, wr);
                std::swap(vi, wi);
                nsets = nout;
            }
            kt_close(c, kt, 0.0, 2.0 * sizeof(T) * (double)m * PB);
            kt = kt_mark(c, KS_LASWP);
            hipLaunchKernelGGL(k_laswp<T>, dim3((n + 255) / 256), dim3(256), 0, s, Ar, Ai, ld, n, j0, PB, c->ipiv);
            kt_close(c, kt, 0.0, 8.0 * sizeof(T) * (double)n * PB);
            const int nright = k0 + NB - (j0 + PB);
            const int nblk = 1 + (m - PB + 255) / 256;
            kt = kt_mark(c, KS_PANEL);
            hipLaunchKernelGGL(k_panel<T>, dim3(nblk), dim3(256), 0, s, Ar, Ai, ld, j0, m, nright, dinfo);
            kt_close(c, kt, 8.0 * (double)m * PB * PB / 2, 4.0 * sizeof(T) * (double)m * PB);
            if (nright > 0) {
                const long o21 = (long)(j0 + PB) * ld + j0, o12 = (long)j0 * ld + j0 + PB;
                const long o22 = (long)(j0 + PB) * ld + j0 + PB;
                cgemm_ctx<T>(c, m - PB, nright, PB, Ar + o21, Ai + o21, ld, Ar + o12, Ai + o12, ld, Ar + o22, Ai + o22,
                         ld, 0);
            }
        }
        const int nr = n - (k0 + NB);
        if (nr > 0) {
            const long o11 = (long)k0 * ld + k0, o12 = (long)k0 * ld + k0 + NB;
            const long o21 = (long)(k0 + NB) * ld + k0, o22 = (long)(k0 + NB) * ld + k0 + NB;
            const int kt = kt_mark(c, KS_TRSM);
            hipLaunchKernelGGL((k_trsm_lower_unit<T, 16>), dim3((nr + 15) / 16), dim3(256), 0, s, NB, Ar + o11,
                               Ai + o11, ld, Ar + o12, Ai + o12, ld, nr);
            kt_close(c, kt, 4.0 * (double)NB * NB * nr, 4.0 * sizeof(T) * (double)NB * nr);
            cgemm_ctx<T>(c, nr, nr, NB, Ar + o21, Ai + o21, ld, Ar + o12, Ai + o12, ld, Ar + o22, Ai + o22, ld, 0);
        }
    }
    return 0;
}

template <typename T>
__global__ void k_convert(const double* __restrict__ a, T* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

}  // namespace

int lu_factor_launch(bem_ctx* c, int precision) {
    const int n = c->npad;
    const size_t nn = (size_t)n * n;
    hipStream_t s = c->stream;
    // candidate buffers: 2 x (n/PB sets x PB ids) ints, 2 x (n x PB) values per plane, + info
    const size_t esz = precision == BEM_F64 ? sizeof(double) : sizeof(float);
    const size_t bytes = 2 * (size_t)n * sizeof(int) + 4 * (size_t)n * PB * esz + 2 * 256 + 64;
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    char* p = static_cast<char*>(scr);
    int* dinfo = reinterpret_cast<int*>(p);
    int* cid = dinfo + 64;
    char* vals = reinterpret_cast<char*>(cid + 2 * (size_t)n);
    BEM_HIP(c, hipMemsetAsync(dinfo, 0, sizeof(int), s));
    if (precision == BEM_F64) {
        BEM_HIP(c, hipMemcpyAsync(c->Vr, c->Zr, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        BEM_HIP(c, hipMemcpyAsync(c->Vi, c->Zi, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        double* cvr = reinterpret_cast<double*>(vals);
        double* cvi = cvr + 2 * (size_t)n * PB;
        factor<double>(c, c->Vr, c->Vi, dinfo, cid, cvr, cvi);
    } else {
        if (!c->has_Z32) {
            const size_t blocks = (nn + 255) / 256;
            hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, s, c->Zr, c->Wr, nn);
            hipLaunchKernelGGL(k_convert<float>, dim3(blocks), dim3(256), 0, s, c->Zi, c->Wi, nn);
        }
        float* cvr = reinterpret_cast<float*>(vals);
        float* cvi = cvr + 2 * (size_t)n * PB;
        factor<float>(c, c->Wr, c->Wi, dinfo, cid, cvr, cvi);
    }
    BEM_HIP(c, hipGetLastError());
    int info = 0;
    BEM_HIP(c, hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    // swap sequence -> row permutation (row i of P Z is row perm[i] of Z)
    std::vector<int> ip(n), perm(n);
    BEM_HIP(c, hipMemcpyAsync(ip.data(), c->ipiv, n * sizeof(int), hipMemcpyDeviceToHost, s));
    BEM_HIP(c, hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++)
        if (ip[i] != i) std::swap(perm[i], perm[ip[i]]);
    BEM_HIP(c, hipMemcpy(c->perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return info > c->n ? 0 : info;  // zero pivots only in the identity padding cannot occur
}
