// solve.hip -- multi-RHS solve with the LU factors and FP64 iterative
// refinement.  Replaces cuSOLVER Zgetrs (solver.cu:51-84, 125-128).
//
// Right-hand sides and solutions are planar FP64, row-major n x ldr (one RHS
// per column), resident on the device.  A solve with the factors of P Z = L U:
//   Y = P B;  Y <- L^{-1} Y;  Y <- U^{-1} Y
// in 256-row blocks: a triangular solve of the diagonal block in LDS
// (trsm.h) followed by a rank-256 GEMM update of the remaining rows on the
// matrix cores (gemm.h), for all right-hand sides at once.
//
// Mixed precision (factors in FP32): X_0 = 0, R_0 = B, and per sweep
//   D = solve32(fp32(P R));  X += D;  R = B - Z X   (FP64 GEMM with the FP64 Z,
//   3M form, K split over batched partial products)
// until max_j ||R_j|| / ||B_j|| <= tol.  Norm reductions use a fixed order.
#include "../bem_internal.h"
#include "gemm.h"
#include "gemm_nt.h"
#include "trsm.h"

namespace {

constexpr int NB = 256;

template <typename T>
__global__ __launch_bounds__(256) void k_gather_rows(int n, int ncol, const int* __restrict__ perm,
                                                     const double* __restrict__ Br, const double* __restrict__ Bi,
                                                     long ldb, T* __restrict__ Yr, T* __restrict__ Yi, long ldy) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * ncol) return;
    const int i = (int)(idx / ncol), j = (int)(idx % ncol);
    const long src = (long)perm[i] * ldb + j;
    Yr[(long)i * ldy + j] = (T)Br[src];
    Yi[(long)i * ldy + j] = (T)Bi[src];
}

template <typename T>
__global__ __launch_bounds__(256) void k_accumulate(long count, const T* __restrict__ Dr, const T* __restrict__ Di,
                                                    double* __restrict__ Xr, double* __restrict__ Xi, int assign) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (assign) {
        Xr[i] = (double)Dr[i];
        Xi[i] = (double)Di[i];
    } else {
        Xr[i] += (double)Dr[i];
        Xi[i] += (double)Di[i];
    }
}

// out[j] = sum_{i < n} |A[i][j]|^2 for j < ncol in two passes with a fixed
// summation order: block b sums rows [256 b, 256 b + 256) of every column
// (64 consecutive columns per wave and row, so the loads are coalesced; the
// four row lanes of a column summed in lane order) into part[b][j]; then
// k_colnorm2_sum adds the blocks' parts in block order.
constexpr int kNormRows = 256;
__global__ __launch_bounds__(256) void k_colnorm2_part(int n, int ncol, const double* __restrict__ Ar,
                                                       const double* __restrict__ Ai, long ld,
                                                       double* __restrict__ part) {
    __shared__ double red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int r0 = blockIdx.x * kNormRows, r1 = min(n, r0 + kNormRows);
    for (int c0 = 0; c0 < ncol; c0 += 64) {
        const int j = c0 + tx;
        double s = 0.0;
        if (j < ncol)
            for (int i = r0 + ty; i < r1; i += 4) {
                const double a = Ar[(long)i * ld + j], b = Ai[(long)i * ld + j];
                s += a * a + b * b;
            }
        red[ty][tx] = s;
        __syncthreads();
        if (ty == 0 && j < ncol) part[(long)blockIdx.x * ncol + j] = ((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx];
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void k_colnorm2_sum(int nblk, int ncol, const double* __restrict__ part,
                                                      double* __restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= ncol) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(long)b * ncol + j];
    out[j] = s;
}
int colnorm_blocks(int n) { return (n + kNormRows - 1) / kNormRows; }
void colnorm2(hipStream_t s, int n, int ncol, const double* Ar, const double* Ai, long ld, double* part, double* out) {
    const int nb = colnorm_blocks(n);
    hipLaunchKernelGGL(k_colnorm2_part, dim3(nb), dim3(256), 0, s, n, ncol, Ar, Ai, ld, part);
    hipLaunchKernelGGL(k_colnorm2_sum, dim3((ncol + 255) / 256), dim3(256), 0, s, nb, ncol, (const double*)part, out);
}

// Updates of at most this many rows go to the short-K kernel (below it the
// tiled GEMM's launches are load-latency bound at K = 256).
constexpr int kShortRows = 4096;

// Y <- U^{-1} L^{-1} Y for the factors at (Fr, Fi), Y planar n x ncol (ld ldy).
//   forward   W_I = L_II^{-1} Y_I;   Y_{>I} -= L_{>I,I} W_I
//   backward  Y_I = U_II^{-1} W_I;   W_{<I} -= U_{<I,I} Y_I
// The updates are matrix-core GEMMs.  The diagonal blocks: FP32 factors apply
// the explicit inverses of the 256 x 256 diagonal blocks at inv (L blocks,
// then U blocks; each n x 256 per plane) as GEMMs, which refinement corrects;
// FP64 factors are solved by substitution (k_trsm_rhs), which keeps the
// unrefined FP64 solve backward stable as Zgetrs is (inv unused).
// W is a work buffer of Y's shape; the result is left in Y.
//
// FP32 with ncol % 32 == 0 and a work area Tt (2 x ncol x kSuper * 256
// floats): two levels.  The 256-blocks of a superblock of kSuper blocks are
// solved with updates inside the superblock only (short-K kernel), their
// solutions also written transposed to Tt by the diagonal products; then one
// update with K = kSuper * 256 brings the superblock's solution into all
// rows beyond it (k_cupdate_nt, both operands k-contiguous: the factor
// columns as they lie in the matrix, Tt).  A rank-1024 update of M rows takes
// 139 us where four rank-256 updates take 170-190 us (C5, 128 columns).
constexpr int kSuper = 4;  // 8 measured the same per solve at C5
constexpr int kSuperCfg = 4;  // NtCfg4: 32 x 32 tiles, 4 blocks per CU (the solves' narrow products)

template <typename T>
void lu_apply(bem_ctx* c, const T* Fr, const T* Fi, const void* inv, T* Yr, T* Yi, T* Wr, T* Wi, int ncol, long ldy,
              T* Tt = nullptr) {
    const int n = c->npad, nb = n / NB;
    const long ld = n;
    constexpr bool subst = sizeof(T) == 8;
    const T* Lir = static_cast<const T*>(inv);
    const T* Lii = subst ? nullptr : Lir + (size_t)n * NB;
    const T* Uir = subst ? nullptr : Lii + (size_t)n * NB;
    const T* Uii = subst ? nullptr : Uir + (size_t)n * NB;
    // diagonal-block products and updates of at most kShortRows rows: the
    // short-K kernel (latency-bound shapes); longer updates: the tiled GEMM
    auto update = [&](int rows, const T* ar, const T* ai, const T* br, const T* bi, T* cr, T* ci) {
        if (rows <= kShortRows)
            cgemm_short_ctx<T>(c, rows, ncol, NB, ar, ai, ld, br, bi, ldy, cr, ci, ldy, 0);
        else
            cgemm_ctx<T>(c, rows, ncol, NB, ar, ai, ld, br, bi, ldy, cr, ci, ldy, 0);
    };
    if constexpr (!subst) {
        auto al = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
        if (Tt && ncol % 32 == 0 && ldy % 4 == 0 && al(Tt) && al(Yr) && al(Yi) && al(Wr) && al(Wi) && al(Fr) &&
            al(Fi) && cgemm_short_fits<T>(NB, ncol, NB, Lir, Lii, NB)) {
            const long ldt = (long)kSuper * NB;
            T* Ttr = Tt;
            T* Tti = Tt + (size_t)ncol * ldt;
            // rows [r0, r0 + rows) -= F[rows, columns [k0, k0 + K)] * Tt[:, :K]^T
            auto super_update = [&](int r0, int rows, int k0, int K, T* cr, T* ci) {
                const long a = (long)r0 * ld + k0;
                const int kt = kt_mark(c, KS_GEMM32, c->stream);
                cupdate_nt(kSuperCfg, c->stream, rows, ncol, K, Fr + a, Fi + a, ld, Ttr, Tti, ldt,
                           cr + (long)r0 * ldy, ci + (long)r0 * ldy, ldy);
                kt_close(c, kt, 6.0 * (double)rows * ncol * K,
                         2.0 * sizeof(T) * ((double)rows * K + (double)K * ncol + 2.0 * rows * ncol));
            };
            for (int I0 = 0; I0 < nb; I0 += kSuper) {
                const int I1 = std::min(I0 + kSuper, nb);
                for (int I = I0; I < I1; I++) {
                    const long y = (long)I * NB * ldy, io = (long)I * NB * NB;
                    cgemm_short_ctx<T>(c, NB, ncol, NB, Lir + io, Lii + io, NB, Yr + y, Yi + y, ldy, Wr + y, Wi + y,
                                       ldy, 1, Ttr + (I - I0) * NB, Tti + (I - I0) * NB, ldt);
                    if (I + 1 < I1) {
                        const long a = (long)(I + 1) * NB * ld + (long)I * NB, yc = (long)(I + 1) * NB * ldy;
                        update((I1 - I - 1) * NB, Fr + a, Fi + a, Wr + y, Wi + y, Yr + yc, Yi + yc);
                    }
                }
                if (I1 < nb) super_update(I1 * NB, n - I1 * NB, I0 * NB, (I1 - I0) * NB, Yr, Yi);
            }
            for (int I1 = nb; I1 > 0; I1 -= kSuper) {
                const int I0 = std::max(I1 - kSuper, 0);
                for (int I = I1 - 1; I >= I0; I--) {
                    const long y = (long)I * NB * ldy, io = (long)I * NB * NB;
                    cgemm_short_ctx<T>(c, NB, ncol, NB, Uir + io, Uii + io, NB, Wr + y, Wi + y, ldy, Yr + y, Yi + y,
                                       ldy, 1, Ttr + (I - I0) * NB, Tti + (I - I0) * NB, ldt);
                    if (I > I0) {
                        const long a = (long)I0 * NB * ld + (long)I * NB, wc = (long)I0 * NB * ldy;
                        update((I - I0) * NB, Fr + a, Fi + a, Yr + y, Yi + y, Wr + wc, Wi + wc);
                    }
                }
                if (I0 > 0) super_update(0, I0 * NB, I0 * NB, (I1 - I0) * NB, Wr, Wi);
            }
            return;
        }
    }
    for (int I = 0; I < nb; I++) {
        const long y = (long)I * NB * ldy, io = (long)I * NB * NB, dg = (long)I * NB * (ld + 1);
        if constexpr (subst)
            trsm_rhs<T>(c->stream, true, Fr + dg, Fi + dg, ld, Yr + y, Yi + y, ldy, Wr + y, Wi + y, ldy, ncol);
        else
            cgemm_short_ctx<T>(c, NB, ncol, NB, Lir + io, Lii + io, NB, Yr + y, Yi + y, ldy, Wr + y, Wi + y, ldy, 1);
        const int rest = n - (I + 1) * NB;
        if (rest > 0) {
            const long a = (long)(I + 1) * NB * ld + (long)I * NB, yc = (long)(I + 1) * NB * ldy;
            update(rest, Fr + a, Fi + a, Wr + y, Wi + y, Yr + yc, Yi + yc);
        }
    }
    for (int I = nb - 1; I >= 0; I--) {
        const long y = (long)I * NB * ldy, io = (long)I * NB * NB, dg = (long)I * NB * (ld + 1);
        if constexpr (subst)
            trsm_rhs<T>(c->stream, false, Fr + dg, Fi + dg, ld, Wr + y, Wi + y, ldy, Yr + y, Yi + y, ldy, ncol);
        else
            cgemm_short_ctx<T>(c, NB, ncol, NB, Uir + io, Uii + io, NB, Wr + y, Wi + y, ldy, Yr + y, Yi + y, ldy, 1);
        if (I > 0) {
            const long a = (long)I * NB;  // rows 0 .. I*NB-1, columns of block I
            update(I * NB, Fr + a, Fi + a, Yr + y, Yi + y, Wr, Wi);
        }
    }
}

// Floats of lu_apply's two-level work area for ncol columns.
size_t lu_super_floats(int ncol) { return 2 * (size_t)ncol * kSuper * NB; }

double max_ratio(const std::vector<double>& r2, const std::vector<double>& b2, int nrhs) {
    double m = 0.0;
    for (int j = 0; j < nrhs; j++) {
        const double v = b2[j] > 0 ? std::sqrt(r2[j] / b2[j]) : std::sqrt(r2[j]);
        m = std::max(m, v);
    }
    return m;
}

}  // namespace

size_t lu_precond_work_bytes(const bem_ctx* c) {
    const size_t ysz = c->factored == BEM_F32_3M ? sizeof(float) : sizeof(double);
    return 4 * (size_t)c->npad * c->ldr * ysz + lu_super_floats(c->ldr) * sizeof(float);
}

int lu_precond_apply(bem_ctx* c, const double* Sr, const double* Si, double* Dr, double* Di, void* work) {
    const int n = c->npad, ncol = c->ldr;
    const long cnt = (long)n * ncol;
    hipStream_t s = c->stream;
    const dim3 g1((unsigned)((cnt + 255) / 256));
    if (c->factored == BEM_F32_3M) {
        float* Yr = static_cast<float*>(work);
        float* Yi = Yr + cnt;
        float* Qr = Yi + cnt;
        float* Qi = Qr + cnt;
        hipLaunchKernelGGL(k_gather_rows<float>, g1, dim3(256), 0, s, n, ncol, c->perm, Sr, Si, (long)ncol, Yr, Yi,
                           (long)ncol);
        lu_apply<float>(c, c->Wr, c->Wi, c->inv32, Yr, Yi, Qr, Qi, ncol, ncol, Qi + cnt);
        hipLaunchKernelGGL(k_accumulate<float>, g1, dim3(256), 0, s, cnt, Yr, Yi, Dr, Di, 1);
    } else if (c->factored == BEM_F64) {
        double* Yr = static_cast<double*>(work);
        double* Yi = Yr + cnt;
        double* Qr = Yi + cnt;
        double* Qi = Qr + cnt;
        hipLaunchKernelGGL(k_gather_rows<double>, g1, dim3(256), 0, s, n, ncol, c->perm, Sr, Si, (long)ncol, Yr, Yi,
                           (long)ncol);
        lu_apply<double>(c, c->Vr, c->Vi, nullptr, Yr, Yi, Qr, Qi, ncol, ncol);
        hipLaunchKernelGGL(k_accumulate<double>, g1, dim3(256), 0, s, cnt, Yr, Yi, Dr, Di, 1);
    } else {
        return bem_fail(c, BEM_ERR_STATE, "preconditioner: no factorisation");
    }
    BEM_HIP(c, hipGetLastError());
    return 0;
}

int lu_solve_launch(bem_ctx* c, int nrhs, double tol, int maxit, int* iters, double* res0, double* res,
                    double* t_solve_ms, double* t_ir_ms) {
    const int n = c->npad, ldr = c->ldr, ncol = ldr;
    const long cnt = (long)n * ldr;
    hipStream_t s = c->stream;
    const bool f32 = c->factored == BEM_F32_3M;
    const bool can_refine = c->has_Z64;
    // scratch: R (FP64 planar), Y and W (working precision planar), norms
    const size_t ysz = f32 ? sizeof(float) : sizeof(double);
    const size_t npart = (size_t)colnorm_blocks(c->n) * ldr;  // column-norm partial sums
    const size_t bytes = 2 * cnt * sizeof(double) + 4 * cnt * ysz + (2 * (size_t)ldr + npart) * sizeof(double) +
                         256 + (f32 ? lu_super_floats(ncol) * sizeof(float) + 256 : 0);
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    double* Rr = static_cast<double*>(scr);
    double* Ri = Rr + cnt;
    char* yp = reinterpret_cast<char*>(Ri + cnt);
    char* wp = yp + 2 * cnt * ysz;
    double* nb2 = reinterpret_cast<double*>(wp + 2 * cnt * ysz);
    double* nr2 = nb2 + ldr;
    double* npt = nr2 + ldr;
    // the two-level FP32 solve's work area, 256-byte aligned
    float* tt = f32 ? reinterpret_cast<float*>((reinterpret_cast<uintptr_t>(npt + npart) + 255) / 256 * 256) : nullptr;
    hipEvent_t e[4];
    for (auto& x : e) BEM_HIP(c, hipEventCreate(&x));
    const dim3 g1((unsigned)((cnt + 255) / 256));
    const dim3 gr((unsigned)(((long)n * ncol + 255) / 256));
    float ts = 0.0f, ti = 0.0f;
    std::vector<double> b2(ldr), r2(ldr);

    auto solve_into_X = [&](const double* Sr, const double* Si, int assign) {
        if (f32) {
            float* Yr = reinterpret_cast<float*>(yp);
            float* Yi = Yr + cnt;
            float* Wr = reinterpret_cast<float*>(wp);
            float* Wi = Wr + cnt;
            hipLaunchKernelGGL(k_gather_rows<float>, gr, dim3(256), 0, s, n, ncol, c->perm, Sr, Si, (long)ldr, Yr, Yi,
                               (long)ldr);
            lu_apply<float>(c, c->Wr, c->Wi, c->inv32, Yr, Yi, Wr, Wi, ncol, ldr, tt);
            hipLaunchKernelGGL(k_accumulate<float>, g1, dim3(256), 0, s, cnt, Yr, Yi, c->Xr, c->Xi, assign);
        } else {
            double* Yr = reinterpret_cast<double*>(yp);
            double* Yi = Yr + cnt;
            double* Wr = reinterpret_cast<double*>(wp);
            double* Wi = Wr + cnt;
            hipLaunchKernelGGL(k_gather_rows<double>, gr, dim3(256), 0, s, n, ncol, c->perm, Sr, Si, (long)ldr, Yr,
                               Yi, (long)ldr);
            lu_apply<double>(c, c->Vr, c->Vi, nullptr, Yr, Yi, Wr, Wi, ncol, ldr);
            hipLaunchKernelGGL(k_accumulate<double>, g1, dim3(256), 0, s, cnt, Yr, Yi, c->Xr, c->Xi, assign);
        }
    };
    auto residual = [&]() -> int {  // R = B - Z X, returns max relative residual in *res
        BEM_HIP(c, hipMemcpyAsync(Rr, c->Br, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
        BEM_HIP(c, hipMemcpyAsync(Ri, c->Bi, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
        // 3M form: three real products (see k_cgemm, ROLE 2)
        if (int rc = cgemm_splitk_ctx<double>(c, n, ncol, n, c->Zr, c->Zi, n, c->Xr, c->Xi, ldr, Rr, Ri, ldr, 0, 2))
            return rc;
        colnorm2(s, c->n, nrhs, Rr, Ri, (long)ldr, npt, nr2);
        BEM_HIP(c, hipMemcpyAsync(r2.data(), nr2, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        BEM_HIP(c, hipStreamSynchronize(s));
        return 0;
    };

    // plain solve
    BEM_HIP(c, hipEventRecord(e[0], s));
    solve_into_X(c->Br, c->Bi, 1);
    BEM_HIP(c, hipEventRecord(e[1], s));
    BEM_HIP(c, hipEventSynchronize(e[1]));
    BEM_HIP(c, hipEventElapsedTime(&ts, e[0], e[1]));
    *iters = 0;
    *res0 = *res = 0.0;
    int rc = 0;
    if (can_refine && (maxit > 0 || f32)) {
        BEM_HIP(c, hipEventRecord(e[2], s));
        colnorm2(s, c->n, nrhs, c->Br, c->Bi, (long)ldr, npt, nb2);
        BEM_HIP(c, hipMemcpyAsync(b2.data(), nb2, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        if (int e2 = residual()) return e2;
        *res0 = *res = max_ratio(r2, b2, nrhs);
        int it = 0;
        float ts_loop = 0.0f;  // triangular solves inside the sweeps
        while (it < maxit && *res > tol) {
            float t0 = 0.0f;
            BEM_HIP(c, hipEventRecord(e[0], s));
            solve_into_X(Rr, Ri, 0);
            BEM_HIP(c, hipEventRecord(e[1], s));
            if (int e2 = residual()) return e2;
            BEM_HIP(c, hipEventElapsedTime(&t0, e[0], e[1]));
            ts_loop += t0;
            *res = max_ratio(r2, b2, nrhs);
            it++;
        }
        *iters = it;
        BEM_HIP(c, hipEventRecord(e[3], s));
        BEM_HIP(c, hipEventSynchronize(e[3]));
        float tall = 0.0f;
        BEM_HIP(c, hipEventElapsedTime(&tall, e[2], e[3]));
        ts += ts_loop;          // all triangular solves
        ti = tall - ts_loop;    // residual GEMMs, norms, updates
        if (maxit > 0 && *res > tol) rc = BEM_ERR_NOCONV;
    }
    BEM_HIP(c, hipGetLastError());
    for (auto& x : e) (void)hipEventDestroy(x);
    *t_solve_ms = ts;
    *t_ir_ms = std::max(0.0f, ti);
    return rc;
}
