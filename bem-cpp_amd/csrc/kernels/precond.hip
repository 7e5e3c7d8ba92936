// precond.hip -- block-Jacobi / restricted additive Schwarz preconditioner
// on the dense matrix (precond.cu:407-895 build, 1303-1410 apply).
//
// The reference assembles each block's 2B x 2B PMCHWT matrix on the CPU
// from its own near-field quadrature, factors it with an unblocked
// partial-pivoting LU (precond.cu:636-868) and applies the factors with one
// warp per block (precond.cu:1303-1410).  On the dense path the block
// matrices are sub-matrices of Z, which is already on the device, and the
// apply has to serve all right-hand sides of an iteration at once, so here:
//   k_bj_gather  copies rows/columns {rwg_j, N + rwg_j : j in the extended
//                block} of the FP64 Z into a zero-padded n2max x n2max slot
//                per block (one launch for all blocks);
//   k_bj_inv     inverts every block in its own workgroup by Gauss-Jordan
//                elimination with the reference's pivoting: at step k the
//                pivot is the first row of largest |a_ik| (precond.cu:842-851),
//                rows k and p are swapped, a pivot below 1e-30 in magnitude
//                is replaced by 1e-15 (precond.cu:858-859); the row swaps
//                become column swaps of the inverse at the end;
//   apply        gathers each block's extended part of r into a padded slot,
//                multiplies all slots by their inverses in one batched
//                FP64 complex GEMM on the matrix cores, and scatters only the
//                block's own unknowns (restricted additive Schwarz; without
//                overlap plain block-Jacobi).
// The inverse gives the same M^{-1} r as the reference's LU solve up to
// rounding, at three times the factorisation work, and turns the apply into
// one matrix-core product instead of sequential substitutions.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../bem_internal.h"
#include "../host/bem_host.h"
#include "gemm.h"

namespace {

// Block b: own size B[b], extended size Be[b] (own first), RWG list at
// rwg + roff[b], pivots at piv + poff[b]; its matrix (then inverse) is slot b
// of ld x ld (row-major per plane, zero beyond 2 Be).
struct BjView {
    const int *B, *Be, *rwg, *roff, *poff;
};

__device__ __forceinline__ int bj_row(const int* rwg, int Be, int N, int p) {
    return p < Be ? rwg[p] : N + rwg[p - Be];
}

__global__ __launch_bounds__(256) void k_bj_gather(BjView v, int N, int ld, const double* __restrict__ Zr,
                                                   const double* __restrict__ Zi, long ldz, double* __restrict__ Ar,
                                                   double* __restrict__ Ai) {
    const int b = blockIdx.x;
    const int Be = v.Be[b], n2 = 2 * Be;
    const int* rwg = v.rwg + v.roff[b];
    double* ar = Ar + (long)b * ld * ld;
    double* ai = Ai + (long)b * ld * ld;
    for (int p = blockIdx.y; p < n2; p += gridDim.y) {
        const long zr = (long)bj_row(rwg, Be, N, p) * ldz;
        for (int q = threadIdx.x; q < n2; q += blockDim.x) {
            const int zc = bj_row(rwg, Be, N, q);
            ar[(long)p * ld + q] = Zr[zr + zc];
            ai[(long)p * ld + q] = Zi[zr + zc];
        }
    }
}

// In-place Gauss-Jordan inverse of the n2 x n2 block at (Ar, Ai) (row-major,
// leading dimension ld), one workgroup per block (1024 threads: one column
// each for blocks up to 1024 unknowns, and enough waves per CU to cover the
// memory latency of the row sweeps); F is an n2-entry column scratch per
// block, piv the row interchanges.
constexpr int BJ_INV_THREADS = 1024;
__global__ __launch_bounds__(BJ_INV_THREADS) void k_bj_inv(BjView v, int ld, double* __restrict__ Ar, double* __restrict__ Ai,
                                                double* __restrict__ Fr, double* __restrict__ Fi,
                                                int* __restrict__ piv) {
    __shared__ double red_v[BJ_INV_THREADS];
    __shared__ int red_i[BJ_INV_THREADS];
    const int b = blockIdx.x, t = threadIdx.x;
    const int n2 = 2 * v.Be[b];
    double* ar = Ar + (long)b * ld * ld;
    double* ai = Ai + (long)b * ld * ld;
    double* fr = Fr + (long)b * ld;
    double* fi = Fi + (long)b * ld;
    int* pv = piv + v.poff[b];
    for (int k = 0; k < n2; k++) {
        // pivot: first row of largest modulus in column k, rows k..n2-1
        double best = -1.0;
        int bi = k;
        for (int i = k + t; i < n2; i += BJ_INV_THREADS) {
            const double m = hypot(ar[(long)i * ld + k], ai[(long)i * ld + k]);
            if (m > best) {
                best = m;
                bi = i;
            }
        }
        red_v[t] = best;
        red_i[t] = bi;
        __syncthreads();
        for (int s = BJ_INV_THREADS / 2; s > 0; s >>= 1) {
            if (t < s) {
                const double o = red_v[t + s];
                const int oi = red_i[t + s];
                if (o > red_v[t] || (o == red_v[t] && oi < red_i[t])) {
                    red_v[t] = o;
                    red_i[t] = oi;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        if (t == 0) pv[k] = p;
        if (p != k)
            for (int j = t; j < n2; j += BJ_INV_THREADS) {
                const long a = (long)k * ld + j, c = (long)p * ld + j;
                const double xr = ar[a], xi = ai[a];
                ar[a] = ar[c];
                ai[a] = ai[c];
                ar[c] = xr;
                ai[c] = xi;
            }
        __syncthreads();
        double dr = ar[(long)k * ld + k], di = ai[(long)k * ld + k];
        if (hypot(dr, di) < 1e-30) {
            dr = 1e-15;
            di = 0.0;
        }
        const double d2 = dr * dr + di * di, qr = dr / d2, qi = -di / d2;  // 1 / pivot
        // column k (the multipliers) into the scratch
        for (int i = t; i < n2; i += BJ_INV_THREADS) {
            fr[i] = ar[(long)i * ld + k];
            fi[i] = ai[(long)i * ld + k];
        }
        __syncthreads();
        // row k scaled by 1 / pivot; its column-k entry becomes 1 / pivot
        for (int j = t; j < n2; j += BJ_INV_THREADS) {
            const long o = (long)k * ld + j;
            if (j == k) {
                ar[o] = qr;
                ai[o] = qi;
            } else {
                const double xr = ar[o], xi = ai[o];
                ar[o] = xr * qr - xi * qi;
                ai[o] = xr * qi + xi * qr;
            }
        }
        __syncthreads();
        // every other row: a_ij -= f_i a_kj (j != k), a_ik = -f_i / pivot
        for (int j = t; j < n2; j += BJ_INV_THREADS) {
            const double ur = ar[(long)k * ld + j], ui = ai[(long)k * ld + j];
#pragma unroll 4
            for (int i = 0; i < n2; i++) {
                if (i == k) continue;
                const long o = (long)i * ld + j;
                const double gr = fr[i], gi = fi[i];
                if (j == k) {
                    ar[o] = -(gr * ur - gi * ui);
                    ai[o] = -(gr * ui + gi * ur);
                } else {
                    ar[o] -= gr * ur - gi * ui;
                    ai[o] -= gr * ui + gi * ur;
                }
            }
        }
        __syncthreads();
    }
    // row interchanges of the elimination = column interchanges of the inverse, in reverse
    for (int k = n2 - 1; k >= 0; k--) {
        const int p = pv[k];
        if (p != k)
            for (int i = t; i < n2; i += BJ_INV_THREADS) {
                const long a = (long)i * ld + k, c = (long)i * ld + p;
                const double xr = ar[a], xi = ai[a];
                ar[a] = ar[c];
                ai[a] = ai[c];
                ar[c] = xr;
                ai[c] = xi;
            }
        __syncthreads();
    }
}

// Slot b of S (ld x ncol per plane) = the extended block of r (rows beyond
// 2 Be stay zero); one workgroup per block.
__global__ __launch_bounds__(256) void k_bj_gather_vec(BjView v, int N, int ld, int ncol,
                                                       const double* __restrict__ Rr, const double* __restrict__ Ri,
                                                       double* __restrict__ Sr, double* __restrict__ Si) {
    const int b = blockIdx.x;
    const int Be = v.Be[b], n2 = 2 * Be;
    const int* rwg = v.rwg + v.roff[b];
    double* sr = Sr + (long)b * ld * ncol;
    double* si = Si + (long)b * ld * ncol;
    for (long e = threadIdx.x; e < (long)n2 * ncol; e += blockDim.x) {
        const int p = (int)(e / ncol), j = (int)(e % ncol);
        const long g = (long)bj_row(rwg, Be, N, p) * ncol + j;
        sr[e] = Rr[g];
        si[e] = Ri[g];
    }
}

// z at the block's own unknowns = rows {0..B-1, Be..Be+B-1} of slot b of Y.
__global__ __launch_bounds__(256) void k_bj_scatter(BjView v, int N, int ld, int ncol, const double* __restrict__ Yr,
                                                    const double* __restrict__ Yi, double* __restrict__ Zr,
                                                    double* __restrict__ Zi) {
    const int b = blockIdx.x;
    const int B = v.B[b], Be = v.Be[b];
    const int* rwg = v.rwg + v.roff[b];
    const double* yr = Yr + (long)b * ld * ncol;
    const double* yi = Yi + (long)b * ld * ncol;
    for (long e = threadIdx.x; e < 2L * B * ncol; e += blockDim.x) {
        const int q = (int)(e / ncol), j = (int)(e % ncol);
        const int p = q < B ? q : Be + (q - B);
        const long g = (long)bj_row(rwg, Be, N, p) * ncol + j;
        Zr[g] = yr[(long)p * ncol + j];
        Zi[g] = yi[(long)p * ncol + j];
    }
}

// rows [n, npad) of the padded system: the identity
__global__ __launch_bounds__(256) void k_copy_tail(long from, long count, const double* __restrict__ Rr,
                                                   const double* __restrict__ Ri, double* __restrict__ Zr,
                                                   double* __restrict__ Zi) {
    const long e = from + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    Zr[e] = Rr[e];
    Zi[e] = Ri[e];
}

template <typename E>
int upload(bem_ctx* c, E*& dst, const std::vector<E>& src, const char* what) {
    if (hipMalloc((void**)&dst, std::max<size_t>(1, src.size()) * sizeof(E)) != hipSuccess) {
        dst = nullptr;
        return bem_fail(c, BEM_ERR_NOMEM, std::string("block_jacobi: ") + what);
    }
    BEM_HIP(c, hipMemcpy((void*)dst, src.data(), src.size() * sizeof(E), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

void bj_free(bem_ctx* c) {
    BlockJacobi& J = c->bj;
    for (void* p : {(void*)J.rwg, (void*)J.B, (void*)J.Be, (void*)J.roff, (void*)J.poff, (void*)J.Ar, (void*)J.Ai,
                    (void*)J.piv})
        if (p) (void)hipFree(p);
    J = BlockJacobi{};
}

int bj_build_launch(bem_ctx* c, double radius_mult, int max_block_rwg, int overlap_layers, double* t_ms) {
    Timer tm;
    const int N = c->rwg.N;
    const BlockPartition P = block_partition(c->mesh, c->rwg, c->quad, radius_mult, max_block_rwg, overlap_layers);
    const int nb = (int)P.own.size();
    if (nb == 0) return bem_fail(c, BEM_ERR_MESH, "block_jacobi: no unknowns");
    std::vector<int> B(nb), Be(nb), roff(nb), poff(nb), rwg;
    long piv_total = 0;
    int n2max = 0;
    for (int b = 0; b < nb; b++) {
        B[b] = (int)P.own[b].size();
        Be[b] = (int)P.ext[b].size();
        if (B[b] < 1 || Be[b] < B[b]) return bem_fail(c, BEM_ERR_STATE, "block_jacobi: inconsistent partition");
        roff[b] = (int)rwg.size();
        for (int m : P.ext[b]) {
            if (m < 0 || m >= N) return bem_fail(c, BEM_ERR_STATE, "block_jacobi: RWG index out of range");
            rwg.push_back(m);
        }
        poff[b] = (int)piv_total;
        piv_total += 2L * Be[b];
        n2max = std::max(n2max, 2 * Be[b]);
    }
    const int ld = (n2max + 63) / 64 * 64;  // slot side: a multiple of the GEMM tile
    const size_t slot = (size_t)ld * ld;
    bj_free(c);
    BlockJacobi& J = c->bj;
    int rc = upload(c, J.rwg, rwg, "RWG lists");
    if (!rc) rc = upload(c, J.B, B, "sizes");
    if (!rc) rc = upload(c, J.Be, Be, "sizes");
    if (!rc) rc = upload(c, J.roff, roff, "offsets");
    if (!rc) rc = upload(c, J.poff, poff, "offsets");
    if (rc) {
        bj_free(c);
        return rc;
    }
    if (hipMalloc((void**)&J.Ar, nb * slot * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&J.Ai, nb * slot * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&J.piv, piv_total * sizeof(int)) != hipSuccess) {
        bj_free(c);
        return bem_fail(c, BEM_ERR_NOMEM, "block_jacobi: inverses (" + std::to_string(nb * slot * 16) + " bytes)");
    }
    void* scr = nullptr;  // the Gauss-Jordan column scratch
    if (int e = ctx_scratch(c, 2 * (size_t)nb * ld * sizeof(double), &scr)) {
        bj_free(c);
        return e;
    }
    double* Fr = static_cast<double*>(scr);
    double* Fi = Fr + (size_t)nb * ld;
    J.nb = nb;
    J.n2max = n2max;
    J.ld = ld;
    J.bytes = nb * slot * 16;
    J.N = N;
    J.own = P.own;
    J.ext = P.ext;
    const BjView v{J.B, J.Be, J.rwg, J.roff, J.poff};
    hipStream_t s = c->stream;
    BEM_HIP(c, hipMemsetAsync(J.Ar, 0, nb * slot * sizeof(double), s));
    BEM_HIP(c, hipMemsetAsync(J.Ai, 0, nb * slot * sizeof(double), s));
    hipLaunchKernelGGL(k_bj_gather, dim3(nb, 64), dim3(256), 0, s, v, N, ld, c->Zr, c->Zi, (long)c->npad, J.Ar,
                       J.Ai);
    hipLaunchKernelGGL(k_bj_inv, dim3(nb), dim3(BJ_INV_THREADS), 0, s, v, ld, J.Ar, J.Ai, Fr, Fi, J.piv);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(s));
    J.valid = true;
    *t_ms = tm.elapsed_ms();
    return nb;
}

size_t bj_work_bytes(const bem_ctx* c) {
    return 4 * (size_t)c->bj.nb * c->bj.ld * c->ldr * sizeof(double);
}

int bj_apply(bem_ctx* c, const double* Rr, const double* Ri, double* Zr, double* Zi, void* work) {
    BlockJacobi& J = c->bj;
    if (!J.valid || J.N * 2 != c->n) return bem_fail(c, BEM_ERR_STATE, "block_jacobi: not built for this system");
    const int ncol = c->ldr, ld = J.ld, nb = J.nb;
    const size_t vs = (size_t)nb * ld * ncol;
    double* Sr = static_cast<double*>(work);
    double* Si = Sr + vs;
    double* Yr = Si + vs;
    double* Yi = Yr + vs;
    hipStream_t s = c->stream;
    const BjView v{J.B, J.Be, J.rwg, J.roff, J.poff};
    BEM_HIP(c, hipMemsetAsync(Sr, 0, 2 * vs * sizeof(double), s));  // padding rows stay zero
    hipLaunchKernelGGL(k_bj_gather_vec, dim3(nb), dim3(256), 0, s, v, J.N, ld, ncol, Rr, Ri, Sr, Si);
    // Y_b = inv_b S_b for every block: one batched FP64 complex GEMM (64 x 64 tile)
    cgemm_any<double>(1, s, ld, ncol, ld, J.Ar, J.Ai, ld, Sr, Si, ncol, Yr, Yi, ncol, 1, nb, (long)ld * ld,
                      (long)ld * ncol, (long)ld * ncol);
    hipLaunchKernelGGL(k_bj_scatter, dim3(nb), dim3(256), 0, s, v, J.N, ld, ncol, Yr, Yi, Zr, Zi);
    const long from = (long)c->n * ncol, count = (long)c->npad * ncol;
    if (count > from)
        hipLaunchKernelGGL(k_copy_tail, dim3((unsigned)((count - from + 255) / 256)), dim3(256), 0, s, from, count,
                           Rr, Ri, Zr, Zi);
    BEM_HIP(c, hipGetLastError());
    return 0;
}
