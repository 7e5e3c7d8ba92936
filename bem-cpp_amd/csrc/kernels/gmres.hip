// gmres.hip -- restarted flexible GMRES on the FP64 matrix, preconditioned on
// the right by the LU factors, for all right-hand sides at once.
//
// Counterpart of the reference's iterative bridge for the dense operator:
// gmres_solve_paired (block_gmres.h:52-61; two systems sharing one matvec)
// and solve_gmres_aext (python/bem_core.py:1104-1123; GMRES on Z with the LU
// of the exterior-only matrix A_ext = [[eta L_ext, -K_ext], [K_ext,
// L_ext/eta]] as preconditioner, python/bem_core.py:1071-1101).  Here every
// right-hand side runs its own Krylov process in lock step with the others:
// one matvec per iteration is one FP64 complex GEMM Z V (n x n x nrhs, 3M
// form, split K), one preconditioner application is one multi-RHS solve with
// the factors on the device (FP32 factors on the matrix cores, or FP64), or
// the block-Jacobi / RAS preconditioner of precond.hip, or none.
// The factors may belong to an earlier matrix: with the factors of Z at
// m = 1 (which is 2 A_ext when eta_ext = 1) this is the reference's A_ext
// preconditioner, reused across refractive indices; with the factors of Z
// itself it is GMRES-based iterative refinement (FGMRES-IR), which converges
// where plain refinement stalls.
//
// Per right-hand side j and cycle (x_0 given, r = b - Z x_0, beta = |r|):
//   v_1 = r / beta;  for i = 1..m:  z_i = M^{-1} v_i;  w = Z z_i;
//   classical Gram-Schmidt twice:  h = V^H w, w -= V h (repeated, summed);
//   h_{i+1,i} = |w|;  v_{i+1} = w / h_{i+1,i};  Givens rotations on the host
//   give the least-squares residual |g_{i+1}|;
//   the right-hand side stops at |g_{i+1}| <= tol |b| (its own m_j);
//   x += sum_{i <= m_j} z_i y_i  with R y = g (back substitution, host).
// The true residual b - Z x (FP64) decides convergence and restarts.  The
// preconditioner is applied in its own rounding, hence the flexible form
// (the z_i are kept).  All device reductions run in a fixed order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "../bem_internal.h"
#include "gemm.h"

namespace {

constexpr int GS_ROWS = 256;  // rows per partial sum of the Gram-Schmidt products

// P[(i * nchunk + chunk) * ncol + j] = sum over the chunk's rows k of
// conj(V_i[k][j]) w[k][j]; grid (nchunk, m), one column per thread.
__global__ __launch_bounds__(256) void k_gs_partial(int n, int ncol, const double* __restrict__ Vr,
                                                    const double* __restrict__ Vi, long vstride,
                                                    const double* __restrict__ wr, const double* __restrict__ wi,
                                                    double* __restrict__ Pr, double* __restrict__ Pi, int nchunk) {
    const int chunk = blockIdx.x, i = blockIdx.y;
    const int k0 = chunk * GS_ROWS, k1 = min(n, k0 + GS_ROWS);
    const double* vr = Vr + (long)i * vstride;
    const double* vi = Vi + (long)i * vstride;
    for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int k = k0; k < k1; k++) {
            const long o = (long)k * ncol + j;
            const double a = vr[o], b = vi[o], c = wr[o], d = wi[o];
            sr += a * c + b * d;  // conj(a + ib) (c + id)
            si += a * d - b * c;
        }
        const long p = ((long)i * nchunk + chunk) * ncol + j;
        Pr[p] = sr;
        Pi[p] = si;
    }
}

// H[i * ncol + j] = sum over chunks (in order) of the partial products.
__global__ __launch_bounds__(256) void k_gs_final(int m, int ncol, int nchunk, const double* __restrict__ Pr,
                                                  const double* __restrict__ Pi, double* __restrict__ Hr,
                                                  double* __restrict__ Hi) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * ncol) return;
    const int i = idx / ncol, j = idx % ncol;
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nchunk; ch++) {
        const long p = ((long)i * nchunk + ch) * ncol + j;
        sr += Pr[p];
        si += Pi[p];
    }
    Hr[idx] = sr;
    Hi[idx] = si;
}

// w[k][j] -= sum_{i < m} V_i[k][j] H[i][j]  (complex; one element per thread)
__global__ __launch_bounds__(256) void k_gs_update(long count, int ncol, int m, const double* __restrict__ Vr,
                                                   const double* __restrict__ Vi, long vstride,
                                                   const double* __restrict__ Hr, const double* __restrict__ Hi,
                                                   double* __restrict__ wr, double* __restrict__ wi) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= count) return;
    const int j = (int)(o % ncol);
    double ar = wr[o], ai = wi[o];
    for (int i = 0; i < m; i++) {
        const double vr = Vr[(long)i * vstride + o], vi = Vi[(long)i * vstride + o];
        const double hr = Hr[i * ncol + j], hi = Hi[i * ncol + j];
        ar -= vr * hr - vi * hi;
        ai -= vr * hi + vi * hr;
    }
    wr[o] = ar;
    wi[o] = ai;
}

// out[j] = sum_k |w[k][j]|^2 over k < n, one block per column, fixed order.
__global__ __launch_bounds__(256) void k_colnorm2_ld(int n, int ncol, const double* __restrict__ wr,
                                                     const double* __restrict__ wi, double* __restrict__ out) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double a = wr[(long)k * ncol + j], b = wi[(long)k * ncol + j];
        s += a * a + b * b;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// dst[k][j] = src[k][j] * s[j] (s real per column)
__global__ __launch_bounds__(256) void k_scale_cols(long count, int ncol, const double* __restrict__ sr,
                                                    const double* __restrict__ si, const double* __restrict__ s,
                                                    double* __restrict__ dr, double* __restrict__ di) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= count) return;
    const double f = s[o % ncol];
    dr[o] = sr[o] * f;
    di[o] = si[o] * f;
}

typedef std::complex<double> cd;

}  // namespace

int gmres_solve_launch(bem_ctx* c, int nrhs, double tol, int restart, int max_iter, int precond, int* iters,
                       double* res0, double* res, double* t_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    const int n = c->npad, ncol = c->ldr, nphys = c->n;
    const long cnt = (long)n * ncol;
    const int m = std::max(1, restart);
    const int nchunk = (n + GS_ROWS - 1) / GS_ROWS;
    hipStream_t s = c->stream;
    // scratch: V (m + 1 vectors), Zb (m vectors), w, r (planar FP64 npad x ldr
    // each), Gram-Schmidt partials and products, column scales, preconditioner work
    const size_t vec = 2 * (size_t)cnt;  // doubles per planar block vector
    const size_t nP = 2 * (size_t)(m + 1) * nchunk * ncol, nH = 2 * (size_t)(m + 1) * ncol;
    const size_t doubles = vec * (2 * (size_t)m + 3) + nP + 2 * nH + 2 * (size_t)ncol;
    const size_t pwork = precond == BEM_PRECOND_LU ? lu_precond_work_bytes(c)
                         : precond == BEM_PRECOND_BLOCK ? bj_work_bytes(c)
                                                        : 0;
    const size_t bytes = doubles * sizeof(double) + pwork + 256;
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    double* base = static_cast<double*>(scr);
    double* V = base;                       // (m + 1) vectors
    double* Zb = V + vec * (m + 1);         // m vectors
    double* w = Zb + vec * m;               // one vector
    double* r = w + vec;                    // one vector
    double* P = r + vec;                    // partials
    double* H = P + nP;                     // Gram-Schmidt products (two passes)
    double* H2 = H + nH;
    double* cs = H2 + nH;                   // column scales
    double* nrm = cs + ncol;                // column norms
    char* pw = reinterpret_cast<char*>(nrm + ncol);
    pw += (256 - reinterpret_cast<uintptr_t>(pw) % 256) % 256;
    void* work = pw;
    auto re = [&](double* v) { return v; };
    auto im = [&](double* v) { return v + cnt; };
    auto Vk = [&](int k) { return V + vec * k; };
    auto Zk = [&](int k) { return Zb + vec * k; };
    const dim3 g1((unsigned)((cnt + 255) / 256));

    std::vector<double> hn2(ncol), bn2(ncol), rn2(ncol), scale(ncol);
    auto colnorm2 = [&](const double* vr, const double* vi, std::vector<double>& out) -> int {
        hipLaunchKernelGGL(k_colnorm2_ld, dim3(ncol), dim3(256), 0, s, nphys, ncol, vr, vi, nrm);
        BEM_HIP(c, hipMemcpyAsync(out.data(), nrm, ncol * sizeof(double), hipMemcpyDeviceToHost, s));
        BEM_HIP(c, hipStreamSynchronize(s));
        return 0;
    };
    // r = B - Z X (FP64, 3M form, split K)
    auto residual = [&]() -> int {
        BEM_HIP(c, hipMemcpyAsync(re(r), c->Br, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
        BEM_HIP(c, hipMemcpyAsync(im(r), c->Bi, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
        return cgemm_splitk_ctx<double>(c, n, ncol, n, c->Zr, c->Zi, n, c->Xr, c->Xi, ncol, re(r), im(r), ncol, 0,
                                        2);
    };
    auto max_rel = [&](const std::vector<double>& r2) {
        double mx = 0.0;
        for (int j = 0; j < nrhs; j++) mx = std::max(mx, bn2[j] > 0 ? std::sqrt(r2[j] / bn2[j]) : 0.0);
        return mx;
    };
    // v_dst = src * scale[j] per column
    auto scale_into = [&](double* src, double* dst) -> int {
        BEM_HIP(c, hipMemcpyAsync(cs, scale.data(), ncol * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_scale_cols, g1, dim3(256), 0, s, cnt, ncol, re(src), im(src), cs, re(dst), im(dst));
        return 0;
    };
    // h (k vectors) = V_{0..k-1}^H w, into Hbuf
    auto gs_products = [&](int k, double* Hbuf) {
        hipLaunchKernelGGL(k_gs_partial, dim3(nchunk, k), dim3(256), 0, s, n, ncol, re(V), im(V), (long)vec, re(w),
                           im(w), P, P + nP / 2, nchunk);
        hipLaunchKernelGGL(k_gs_final, dim3((k * ncol + 255) / 256), dim3(256), 0, s, k, ncol, nchunk, P, P + nP / 2,
                           Hbuf, Hbuf + nH / 2);
        hipLaunchKernelGGL(k_gs_update, g1, dim3(256), 0, s, cnt, ncol, k, re(V), im(V), (long)vec, Hbuf,
                           Hbuf + nH / 2, re(w), im(w));
    };

    BEM_HIP(c, hipMemsetAsync(c->Xr, 0, cnt * sizeof(double), s));
    BEM_HIP(c, hipMemsetAsync(c->Xi, 0, cnt * sizeof(double), s));
    if (int rc = colnorm2(c->Br, c->Bi, bn2)) return rc;
    *res0 = 1.0;  // x_0 = 0
    *res = nrhs > 0 && max_rel(bn2) > 0 ? 1.0 : 0.0;
    int total = 0;
    // host Hessenberg data per right-hand side: column-major (m + 1) x m
    std::vector<cd> Hh((size_t)(m + 1) * m * ncol), g((size_t)(m + 1) * ncol), sn((size_t)m * ncol);
    std::vector<double> cn((size_t)m * ncol);
    std::vector<cd> hcol((size_t)(m + 1) * ncol);
    std::vector<double> hbr((size_t)(m + 1) * ncol), hbi((size_t)(m + 1) * ncol);
    std::vector<int> mj(ncol), active(ncol);
    bool first = true, checked = false;
    while (total < max_iter) {
        // r = b - Z x (x = 0 on the first cycle: r = b)
        if (first) {
            BEM_HIP(c, hipMemcpyAsync(re(r), c->Br, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
            BEM_HIP(c, hipMemcpyAsync(im(r), c->Bi, cnt * sizeof(double), hipMemcpyDeviceToDevice, s));
            rn2 = bn2;
        } else {
            if (int rc = residual()) return rc;
            if (int rc = colnorm2(re(r), im(r), rn2)) return rc;
            *res = max_rel(rn2);
            if (*res <= tol) {
                checked = true;
                break;
            }
        }
        first = false;
        int nact = 0;
        for (int j = 0; j < ncol; j++) {
            const double beta = std::sqrt(rn2[j]);
            active[j] = j < nrhs && bn2[j] > 0 && beta > tol * std::sqrt(bn2[j]);
            nact += active[j];
            mj[j] = 0;
            scale[j] = active[j] ? 1.0 / beta : 0.0;
            g[(size_t)j * (m + 1)] = active[j] ? cd(beta, 0.0) : cd(0.0, 0.0);
        }
        if (nact == 0) break;
        if (int rc = scale_into(r, Vk(0))) return rc;
        int it = 0;
        for (; it < m && total < max_iter; it++, total++) {
            int prc = 0;
            if (precond == BEM_PRECOND_LU)
                prc = lu_precond_apply(c, re(Vk(it)), im(Vk(it)), re(Zk(it)), im(Zk(it)), work);
            else if (precond == BEM_PRECOND_BLOCK)
                prc = bj_apply(c, re(Vk(it)), im(Vk(it)), re(Zk(it)), im(Zk(it)), work);
            else
                BEM_HIP(c, hipMemcpyAsync(Zk(it), Vk(it), vec * sizeof(double), hipMemcpyDeviceToDevice, s));
            if (prc) return prc;
            if (int rc = cgemm_splitk_ctx<double>(c, n, ncol, n, c->Zr, c->Zi, n, re(Zk(it)), im(Zk(it)), ncol, re(w),
                                                  im(w), ncol, 1, 2))
                return rc;
            // classical Gram-Schmidt, twice
            gs_products(it + 1, H);
            gs_products(it + 1, H2);
            BEM_HIP(c, hipMemcpyAsync(hbr.data(), H, (size_t)(it + 1) * ncol * sizeof(double), hipMemcpyDeviceToHost,
                                      s));
            BEM_HIP(c, hipMemcpyAsync(hbi.data(), H + nH / 2, (size_t)(it + 1) * ncol * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
            std::vector<double> h2r((size_t)(it + 1) * ncol), h2i((size_t)(it + 1) * ncol);
            BEM_HIP(c, hipMemcpyAsync(h2r.data(), H2, h2r.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            BEM_HIP(c, hipMemcpyAsync(h2i.data(), H2 + nH / 2, h2i.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            if (int rc = colnorm2(re(w), im(w), hn2)) return rc;
            int still = 0;
            for (int j = 0; j < ncol; j++) {
                if (!active[j]) {
                    scale[j] = 0.0;
                    continue;
                }
                cd* Hj = &Hh[(size_t)j * (m + 1) * m + (size_t)it * (m + 1)];  // column it
                for (int i = 0; i <= it; i++)
                    Hj[i] = cd(hbr[(size_t)i * ncol + j] + h2r[(size_t)i * ncol + j],
                               hbi[(size_t)i * ncol + j] + h2i[(size_t)i * ncol + j]);
                const double hnext = std::sqrt(hn2[j]);
                // earlier rotations, then a new one that zeroes h_{it+1,it}
                for (int i = 0; i < it; i++) {
                    const double ci = cn[(size_t)j * m + i];
                    const cd si = sn[(size_t)j * m + i];
                    const cd a = Hj[i], b = Hj[i + 1];
                    Hj[i] = ci * a + si * b;
                    Hj[i + 1] = -std::conj(si) * a + ci * b;
                }
                const cd a = Hj[it];
                const double aa = std::abs(a), rho = std::hypot(aa, hnext);
                double cr;
                cd sr;
                if (rho == 0.0) {
                    cr = 1.0;
                    sr = 0.0;
                } else if (aa == 0.0) {
                    cr = 0.0;
                    sr = 1.0;
                } else {
                    cr = aa / rho;
                    sr = (a / aa) * hnext / rho;
                }
                cn[(size_t)j * m + it] = cr;
                sn[(size_t)j * m + it] = sr;
                Hj[it] = cr * a + sr * hnext;
                cd* gj = &g[(size_t)j * (m + 1)];
                gj[it + 1] = -std::conj(sr) * gj[it];
                gj[it] = cr * gj[it];
                mj[j] = it + 1;
                const bool done = std::abs(gj[it + 1]) <= tol * std::sqrt(bn2[j]) || hnext == 0.0;
                if (done) {
                    active[j] = 0;
                    scale[j] = 0.0;
                } else {
                    scale[j] = 1.0 / hnext;
                    still++;
                }
            }
            if (it + 1 < m) {
                if (int rc = scale_into(w, Vk(it + 1))) return rc;
            }
            if (still == 0) {
                it++;
                total++;
                break;
            }
        }
        // y per right-hand side (R y = g), then x += sum_i z_i y_i
        const int mmax = *std::max_element(mj.begin(), mj.end());
        if (mmax == 0) break;
        std::vector<double> yr((size_t)mmax * ncol, 0.0), yi((size_t)mmax * ncol, 0.0);
        for (int j = 0; j < ncol; j++) {
            const int k = mj[j];
            if (k == 0) continue;
            const cd* Hj = &Hh[(size_t)j * (m + 1) * m];
            const cd* gj = &g[(size_t)j * (m + 1)];
            std::vector<cd> y(k);
            for (int i = k - 1; i >= 0; i--) {
                cd acc = gj[i];
                for (int l = i + 1; l < k; l++) acc -= Hj[(size_t)l * (m + 1) + i] * y[l];
                const cd d = Hj[(size_t)i * (m + 1) + i];
                y[i] = d != cd(0.0, 0.0) ? acc / d : cd(0.0, 0.0);
            }
            for (int i = 0; i < k; i++) {
                // the update subtracts, so store -y
                yr[(size_t)i * ncol + j] = -y[i].real();
                yi[(size_t)i * ncol + j] = -y[i].imag();
            }
        }
        BEM_HIP(c, hipMemcpyAsync(H, yr.data(), yr.size() * sizeof(double), hipMemcpyHostToDevice, s));
        BEM_HIP(c, hipMemcpyAsync(H + nH / 2, yi.data(), yi.size() * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gs_update, g1, dim3(256), 0, s, cnt, ncol, mmax, re(Zb), im(Zb), (long)vec, H,
                           H + nH / 2, c->Xr, c->Xi);
        BEM_HIP(c, hipStreamSynchronize(s));
    }
    // final true residual (unless the loop ended on one)
    if (!checked) {
        if (int rc = residual()) return rc;
        if (int rc = colnorm2(re(r), im(r), rn2)) return rc;
        *res = max_rel(rn2);
    }
    *iters = total;
    BEM_HIP(c, hipGetLastError());
    *t_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return *res <= tol ? 0 : BEM_ERR_NOCONV;
}
