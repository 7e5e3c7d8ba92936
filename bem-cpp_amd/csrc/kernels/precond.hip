// precond.hip -- block-Jacobi / restricted additive Schwarz preconditioner
// on the dense matrix (precond.cu:407-895 build, 1303-1410 apply).
//
// The reference assembles each block's 2B x 2B PMCHWT matrix on the CPU
// from its own near-field quadrature and factors it with an unblocked
// partial-pivoting LU (precond.cu:636-868); on the dense path the block
// matrices are sub-matrices of Z, which is already on the device, so here:
//   k_bj_gather  copies rows/columns {rwg_j, N + rwg_j : j in the extended
//                block} of the FP64 Z into a row-major block (one launch for
//                all blocks);
//   k_bj_lu      factors every block in its own workgroup: at step k the
//                pivot is the first row of largest |a_ik| (precond.cu:842-851),
//                rows k and p are swapped whole, a pivot below 1e-30 in
//                magnitude is replaced by 1e-15 (precond.cu:858-859; the
//                replacement is also stored, so the solve divides by it),
//                then multipliers and the rank-1 update;
//   k_bj_apply   per block and right-hand side: gather the extended block of
//                r, the row interchanges, unit-lower and upper substitution,
//                and scatter of the block's own unknowns only (restricted
//                additive Schwarz; without overlap plain block-Jacobi).
// One thread owns one right-hand-side column through the substitutions, so
// every factor entry is read once per block for all right-hand sides.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../bem_internal.h"
#include "../host/bem_host.h"

namespace {

// Block b: own size B[b], extended size Be[b] (own first), RWG list at
// rwg + roff[b], factors at lu + loff[b] (Be2 x Be2 row-major per plane,
// Be2 = 2 Be), pivots at piv + poff[b].
struct BjView {
    const int *B, *Be, *rwg, *roff, *poff;
    const long* loff;
};

__device__ __forceinline__ int bj_row(const int* rwg, int Be, int N, int p) {
    return p < Be ? rwg[p] : N + rwg[p - Be];
}

__global__ __launch_bounds__(256) void k_bj_gather(BjView v, int N, const double* __restrict__ Zr,
                                                   const double* __restrict__ Zi, long ldz, double* __restrict__ Lr,
                                                   double* __restrict__ Li) {
    const int b = blockIdx.x;
    const int Be = v.Be[b], n2 = 2 * Be;
    const int* rwg = v.rwg + v.roff[b];
    double* ar = Lr + v.loff[b];
    double* ai = Li + v.loff[b];
    for (int p = blockIdx.y; p < n2; p += gridDim.y) {
        const long zr = (long)bj_row(rwg, Be, N, p) * ldz;
        for (int q = threadIdx.x; q < n2; q += blockDim.x) {
            const int zc = bj_row(rwg, Be, N, q);
            ar[(long)p * n2 + q] = Zr[zr + zc];
            ai[(long)p * n2 + q] = Zi[zr + zc];
        }
    }
}

__global__ __launch_bounds__(256) void k_bj_lu(BjView v, double* __restrict__ Lr, double* __restrict__ Li,
                                               int* __restrict__ piv) {
    __shared__ double red_v[256];
    __shared__ int red_i[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const int n2 = 2 * v.Be[b];
    double* ar = Lr + v.loff[b];
    double* ai = Li + v.loff[b];
    int* pv = piv + v.poff[b];
    for (int k = 0; k < n2; k++) {
        // pivot: first row of largest modulus in column k
        double best = -1.0;
        int bi = k;
        for (int i = k + t; i < n2; i += 256) {
            const double m = hypot(ar[(long)i * n2 + k], ai[(long)i * n2 + k]);
            if (m > best) {
                best = m;
                bi = i;
            }
        }
        red_v[t] = best;
        red_i[t] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
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
            for (int j = t; j < n2; j += 256) {
                const long a = (long)k * n2 + j, c = (long)p * n2 + j;
                const double xr = ar[a], xi = ai[a];
                ar[a] = ar[c];
                ai[a] = ai[c];
                ar[c] = xr;
                ai[c] = xi;
            }
        __syncthreads();
        double dr = ar[(long)k * n2 + k], di = ai[(long)k * n2 + k];
        if (hypot(dr, di) < 1e-30) {
            dr = 1e-15;
            di = 0.0;
        }
        __syncthreads();  // every thread has read the pivot before it is replaced
        if (t == 0) {
            ar[(long)k * n2 + k] = dr;
            ai[(long)k * n2 + k] = di;
        }
        const double d2 = dr * dr + di * di;
        for (int i = k + 1 + t; i < n2; i += 256) {  // multipliers a_ik / a_kk
            const long o = (long)i * n2 + k;
            const double xr = ar[o], xi = ai[o];
            ar[o] = (xr * dr + xi * di) / d2;
            ai[o] = (xi * dr - xr * di) / d2;
        }
        __syncthreads();
        for (int j = k + 1 + t; j < n2; j += 256) {  // rank-1 update, one column per thread
            const double ur = ar[(long)k * n2 + j], ui = ai[(long)k * n2 + j];
            for (int i = k + 1; i < n2; i++) {
                const long o = (long)i * n2 + j, l = (long)i * n2 + k;
                const double lr = ar[l], li = ai[l];
                ar[o] -= lr * ur - li * ui;
                ai[o] -= lr * ui + li * ur;
            }
        }
        __syncthreads();
    }
}

// z (own unknowns of each block) = factors^{-1} r (extended block), all
// right-hand sides; R, Z planar FP64 npad x ld; W per-block work at woff.
__global__ __launch_bounds__(256) void k_bj_apply(BjView v, int N, int ncol, const double* __restrict__ Lr,
                                                  const double* __restrict__ Li, const int* __restrict__ piv,
                                                  const double* __restrict__ Rr, const double* __restrict__ Ri,
                                                  double* __restrict__ Zr, double* __restrict__ Zi,
                                                  double* __restrict__ Wr, double* __restrict__ Wi,
                                                  const long* __restrict__ woff) {
    const int b = blockIdx.x;
    const int B = v.B[b], Be = v.Be[b], n2 = 2 * Be;
    const int* rwg = v.rwg + v.roff[b];
    const double* ar = Lr + v.loff[b];
    const double* ai = Li + v.loff[b];
    const int* pv = piv + v.poff[b];
    double* yr = Wr + woff[b];
    double* yi = Wi + woff[b];
    for (long e = threadIdx.x; e < (long)n2 * ncol; e += blockDim.x) {
        const int p = (int)(e / ncol), j = (int)(e % ncol);
        const long g = (long)bj_row(rwg, Be, N, p) * ncol + j;
        yr[e] = Rr[g];
        yi[e] = Ri[g];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
        for (int k = 0; k < n2; k++) {
            const int p = pv[k];
            if (p != k) {
                const long a = (long)k * ncol + j, c = (long)p * ncol + j;
                const double xr = yr[a], xi = yi[a];
                yr[a] = yr[c];
                yi[a] = yi[c];
                yr[c] = xr;
                yi[c] = xi;
            }
        }
        for (int i = 1; i < n2; i++) {  // unit lower
            double sr = 0.0, si = 0.0;
            for (int l = 0; l < i; l++) {
                const double lr = ar[(long)i * n2 + l], li = ai[(long)i * n2 + l];
                const double xr = yr[(long)l * ncol + j], xi = yi[(long)l * ncol + j];
                sr += lr * xr - li * xi;
                si += lr * xi + li * xr;
            }
            yr[(long)i * ncol + j] -= sr;
            yi[(long)i * ncol + j] -= si;
        }
        for (int i = n2 - 1; i >= 0; i--) {  // upper
            double sr = yr[(long)i * ncol + j], si = yi[(long)i * ncol + j];
            for (int l = i + 1; l < n2; l++) {
                const double ur = ar[(long)i * n2 + l], ui = ai[(long)i * n2 + l];
                const double xr = yr[(long)l * ncol + j], xi = yi[(long)l * ncol + j];
                sr -= ur * xr - ui * xi;
                si -= ur * xi + ui * xr;
            }
            const double dr = ar[(long)i * n2 + i], di = ai[(long)i * n2 + i], d2 = dr * dr + di * di;
            yr[(long)i * ncol + j] = (sr * dr + si * di) / d2;
            yi[(long)i * ncol + j] = (si * dr - sr * di) / d2;
        }
    }
    __syncthreads();
    // the block's own unknowns only: J parts p < B, M parts Be <= p < Be + B
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
    for (void* p : {(void*)J.rwg, (void*)J.B, (void*)J.Be, (void*)J.roff, (void*)J.poff, (void*)J.loff,
                    (void*)J.woff, (void*)J.Lr, (void*)J.Li, (void*)J.piv})
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
    std::vector<long> loff(nb);
    long lu_total = 0, piv_total = 0, w_total = 0;
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
        const long n2 = 2L * Be[b];
        loff[b] = lu_total;
        lu_total += n2 * n2;
        poff[b] = (int)piv_total;
        piv_total += n2;
        w_total += n2;  // rows of substitution work, times the column count at apply time
        n2max = std::max(n2max, (int)n2);
    }
    bj_free(c);
    BlockJacobi& J = c->bj;
    int rc = upload(c, J.rwg, rwg, "RWG lists");
    if (!rc) rc = upload(c, J.B, B, "sizes");
    if (!rc) rc = upload(c, J.Be, Be, "sizes");
    if (!rc) rc = upload(c, J.roff, roff, "offsets");
    if (!rc) rc = upload(c, J.poff, poff, "offsets");
    if (!rc) rc = upload(c, J.loff, loff, "offsets");
    if (!rc) rc = upload(c, J.woff, std::vector<long>(nb, 0), "offsets");
    if (rc) {
        bj_free(c);
        return rc;
    }
    if (hipMalloc((void**)&J.Lr, lu_total * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&J.Li, lu_total * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&J.piv, piv_total * sizeof(int)) != hipSuccess) {
        bj_free(c);
        return bem_fail(c, BEM_ERR_NOMEM, "block_jacobi: factors (" + std::to_string(lu_total * 16) + " bytes)");
    }
    J.nb = nb;
    J.n2max = n2max;
    J.work_rows = w_total;
    J.lu_bytes = (size_t)lu_total * 16;
    J.N = N;
    J.own = P.own;
    J.ext = P.ext;
    const BjView v{J.B, J.Be, J.rwg, J.roff, J.poff, J.loff};
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_bj_gather, dim3(nb, 64), dim3(256), 0, s, v, N, c->Zr, c->Zi, (long)c->npad, J.Lr, J.Li);
    hipLaunchKernelGGL(k_bj_lu, dim3(nb), dim3(256), 0, s, v, J.Lr, J.Li, J.piv);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(s));
    J.valid = true;
    *t_ms = tm.elapsed_ms();
    return nb;
}

size_t bj_work_bytes(const bem_ctx* c) { return 2 * (size_t)c->bj.work_rows * c->ldr * sizeof(double); }

int bj_apply(bem_ctx* c, const double* Rr, const double* Ri, double* Zr, double* Zi, void* work) {
    BlockJacobi& J = c->bj;
    if (!J.valid || J.N * 2 != c->n) return bem_fail(c, BEM_ERR_STATE, "block_jacobi: not built for this system");
    const int ncol = c->ldr;
    double* Wr = static_cast<double*>(work);
    double* Wi = Wr + (size_t)J.work_rows * ncol;
    if (J.woff_ncol != ncol) {  // per-block work offsets for this column count
        std::vector<long> woff(J.nb);
        long acc = 0;
        for (int b = 0; b < J.nb; b++) {
            woff[b] = acc * ncol;
            acc += 2L * (long)J.ext[b].size();
        }
        BEM_HIP(c, hipStreamSynchronize(c->stream));
        BEM_HIP(c, hipMemcpy(J.woff, woff.data(), J.nb * sizeof(long), hipMemcpyHostToDevice));
        J.woff_ncol = ncol;
    }
    const BjView v{J.B, J.Be, J.rwg, J.roff, J.poff, J.loff};
    hipLaunchKernelGGL(k_bj_apply, dim3(J.nb), dim3(256), 0, c->stream, v, J.N, ncol, J.Lr, J.Li, J.piv, Rr, Ri, Zr,
                       Zi, Wr, Wi, J.woff);
    const long from = (long)c->n * ncol, count = (long)c->npad * ncol;
    if (count > from)
        hipLaunchKernelGGL(k_copy_tail, dim3((unsigned)((count - from + 255) / 256)), dim3(256), 0, c->stream, from,
                           count, Rr, Ri, Zr, Zi);
    BEM_HIP(c, hipGetLastError());
    return 0;
}
