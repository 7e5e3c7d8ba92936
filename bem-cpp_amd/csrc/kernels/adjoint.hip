// adjoint.hip -- gradient of a real loss of the orientation-averaged Mueller
// matrix with respect to p = (Re m, Im m, ka), for k_ext = ka, k_int = ka m,
// eta_ext = 1, eta_int = 1/|m| (main.cpp:192-196).  The reference has no
// derivatives; this is the adjoint of its dense path.
//
// With X the solutions of Z X = B and g = dL/dX (holomorphic derivative, so
// dL = 2 Re sum g dX for the real L):
//   dL/dp = 2 Re[ sum_c lambda_c^T dB_c/dp ] - 2 Re[ <dZ/dp, W> ] + dL/dp|_X
//   Z^T lambda_c = g_c,   W = Lambda X^T,   <A, W> = sum_{mn} A_mn W_mn.
// Z^T = S Z S with S = diag(I, -I), so lambda = S Z^{-1} S g reuses the
// factorisation (FP32 factors + FP64 refinement).  dZ/dp is never stored: two
// assembly passes per parameter write half-pair records of dZ/dp (one from
// the moments of G with the k-derivatives of the prefactors and of eta_i, one
// from the moments of dG/dk) and k_contract reduces them against W.  Only ka
// reaches B (the plane-wave phases) and the far field (phases, the -ik
// prefactors and the 1/|k|^2 normalisation).
#include <cmath>

#include <cstdlib>

#include "../bem_internal.h"
#include "gemm.h"

namespace {

__global__ __launch_bounds__(256) void k_negate_lower(long lo, long hi, double* __restrict__ a,
                                                      double* __restrict__ b) {
    const long i = lo + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    a[i] = -a[i];
    b[i] = -b[i];
}

// T[j][i] = A[i][j] for the planar n x m matrix A (ld lda) into T (ld ldt).
__global__ __launch_bounds__(256) void k_transpose(int n, int m, const double* __restrict__ Ar,
                                                   const double* __restrict__ Ai, long lda, double* __restrict__ Tr,
                                                   double* __restrict__ Ti, long ldt) {
    __shared__ double sr[32][33], si[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: column block of A, by: row block
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int i = by + r, j = bx + tx;
        if (i < n && j < m) {
            sr[r][tx] = Ar[(long)i * lda + j];
            si[r][tx] = Ai[(long)i * lda + j];
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = bx + r, i = by + tx;
        if (i < n && j < m) {
            Tr[(long)j * ldt + i] = sr[tx][r];
            Ti[(long)j * ldt + i] = si[tx][r];
        }
    }
}

// sum_{i < n, j < m} A[i][j] B[i][j] (complex, no conjugate), one block per
// row block, fixed order; part[blockIdx.x] = (re, im).
__global__ __launch_bounds__(256) void k_dot_rows(int n, int m, const double* __restrict__ Ar,
                                                  const double* __restrict__ Ai, const double* __restrict__ Br,
                                                  const double* __restrict__ Bi, long ld, double* __restrict__ part) {
    __shared__ double red[2][256];
    double s0 = 0.0, s1 = 0.0;
    const long total = (long)n * m;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e / m, j = e % m, o = i * ld + j;
        s0 += Ar[o] * Br[o] - Ai[o] * Bi[o];
        s1 += Ar[o] * Bi[o] + Ai[o] * Br[o];
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0];
        part[2 * blockIdx.x + 1] = red[1][0];
    }
}

struct Cx {
    double r, i;
};
inline Cx cmul(Cx a, Cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
inline Cx cscale(Cx a, double s) { return {a.r * s, a.i * s}; }
inline Cx cadd(Cx a, Cx b) { return {a.r + b.r, a.i + b.i}; }
inline Cx cinv(Cx a) {
    const double d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}
inline void cset(double* out, Cx v) {
    out[0] = v.r;
    out[1] = v.i;
}

// Record coefficients of dZ/dp (see RecCoef): pass 0 uses the moments of G, g
// with the derivatives of the prefactors (i k, -4i/k) and of eta_w; pass 1 the
// moments of dG/dk, dg/dk.  kappa[w] = dk_w/dp, deta[w] = d eta_w/dp,
// dieta[w] = d(1/eta_w)/dp.
RecCoef make_coef(int pass, const Cx k[2], const double eta[2], const Cx kappa[2], const double deta[2],
                  const double dieta[2]) {
    RecCoef q{};
    q.deriv = pass;
    const Cx I{0.0, 1.0};
    for (int w = 0; w < 2; w++) {
        const Cx ik = cmul(I, k[w]);                         // i k
        const Cx m4ik = cscale(cmul(I, cinv(k[w])), -4.0);  // -4i/k
        const Cx dik = cmul(I, kappa[w]);                   // d(ik)/dp
        // d(-4i/k)/dp = 4i kappa / k^2
        const Cx dm4ik = cscale(cmul(cmul(I, kappa[w]), cinv(cmul(k[w], k[w]))), 4.0);
        const double ie = 1.0 / eta[w];
        Cx aP, bP, aQ, bQ, cK;
        if (pass == 0) {
            aP = cadd(cscale(ik, deta[w]), cscale(dik, eta[w]));
            bP = cadd(cscale(m4ik, deta[w]), cscale(dm4ik, eta[w]));
            aQ = cadd(cscale(ik, dieta[w]), cscale(dik, ie));
            bQ = cadd(cscale(m4ik, dieta[w]), cscale(dm4ik, ie));
            cK = {0.0, 0.0};
        } else {
            aP = cscale(cmul(kappa[w], ik), eta[w]);
            bP = cscale(cmul(kappa[w], m4ik), eta[w]);
            aQ = cscale(cmul(kappa[w], ik), ie);
            bQ = cscale(cmul(kappa[w], m4ik), ie);
            cK = kappa[w];
        }
        cset(q.aP[w], aP);
        cset(q.bP[w], bP);
        cset(q.aQ[w], aQ);
        cset(q.bQ[w], bQ);
        cset(q.cK[w], cK);
        // self pairs at R = 0: the forward value i Re(k)/(4 pi) changes by
        // i Re(kappa)/(4 pi); the records multiply dG/dk by kappa
        Cx r0{0.0, 0.0};
        if (pass == 1 && (kappa[w].r != 0.0 || kappa[w].i != 0.0))
            r0 = cmul(Cx{0.0, kappa[w].r / (4.0 * M_PI)}, cinv(kappa[w]));
        cset(q.r0[w], r0);
    }
    return q;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int adjoint_launch(bem_ctx* c, const double* dL_dM, double* grad) {
    if (!c->has_Z64) return bem_fail(c, BEM_ERR_STATE, "mueller_grad: needs the FP64 matrix (assemble with Z64)");
    if (!c->has_X || !c->rhs_planewave)
        return bem_fail(c, BEM_ERR_STATE, "mueller_grad: needs plane-wave right-hand sides and their solutions");
    const AsmParams& ap = c->asm_params;
    const double ka = ap.kr[0];
    if (ap.ki[0] != 0.0 || c->last_kext[0] != ka || c->last_kext[1] != 0.0 || c->rhs_k[0] != ka ||
        c->rhs_k[1] != 0.0 || ap.eta[0] != 1.0 || c->last_eta_ext != 1.0 || c->rhs_eta != 1.0 || !(ka > 0.0))
        return bem_fail(c, BEM_ERR_STATE,
                        "mueller_grad: defined for k_ext = ka real, eta_ext = 1 in assembly, RHS and far field");
    const Cx m{ap.kr[1] / ka, ap.ki[1] / ka};
    const double am = std::hypot(m.r, m.i);
    if (std::fabs(ap.eta[1] * am - 1.0) > 1e-12)
        return bem_fail(c, BEM_ERR_STATE, "mueller_grad: needs eta_int = 1/|m| with m = k_int/ka");
    if (2 * c->last_norient > c->nrhs)
        return bem_fail(c, BEM_ERR_STATE, "mueller_grad: the solutions do not cover the orientations");

    const int npad = c->npad, N = c->rwg.N, ldr = c->ldr;
    const long cnt = (long)npad * ldr;
    hipStream_t s = c->stream;
    // work: saved B and X, Lambda, X^T, dB, W, partial sums
    const int nblk_dot = 512;
    const size_t doubles = 2 * cnt * 5 + 2 * (size_t)npad * npad + 2 * (size_t)nblk_dot + 64;
    DevBuf buf;
    if (hipMalloc(&buf.p, doubles * sizeof(double)) != hipSuccess)
        return bem_fail(c, BEM_ERR_NOMEM, "mueller_grad: work area (" + std::to_string(doubles * 8) + " bytes)");
    double* Bsr = static_cast<double*>(buf.p);
    double* Bsi = Bsr + cnt;
    double* Xsr = Bsi + cnt;
    double* Xsi = Xsr + cnt;
    double* Lr = Xsi + cnt;
    double* Li = Lr + cnt;
    double* XTr = Li + cnt;
    double* XTi = XTr + cnt;
    double* dBr = XTi + cnt;
    double* dBi = dBr + cnt;
    double* Wr = dBi + cnt;
    double* Wi = Wr + (size_t)npad * npad;
    double* part = Wi + (size_t)npad * npad;

    BEM_HIP(c, hipMemcpyAsync(Bsr, c->Br, cnt * 8, hipMemcpyDeviceToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(Bsi, c->Bi, cnt * 8, hipMemcpyDeviceToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(Xsr, c->Xr, cnt * 8, hipMemcpyDeviceToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(Xsi, c->Xi, cnt * 8, hipMemcpyDeviceToDevice, s));
    // B and X are overwritten by the adjoint solve below; every exit from here
    // on puts the forward right-hand sides and solutions back (or, if even
    // the copies fail, marks them as gone)
    struct Restore {
        bem_ctx* c;
        const double *br, *bi, *xr, *xi;
        long cnt;
        bool done = false;
        int run() {
            if (done) return 0;
            done = true;
            hipError_t e = hipMemcpyAsync(c->Br, br, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(c->Bi, bi, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(c->Xr, xr, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(c->Xi, xi, cnt * 8, hipMemcpyDeviceToDevice, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // before the save buffers are freed
            if (e != hipSuccess) {
                c->has_B = c->has_X = false;
                c->last_norient = 0;
                return bem_fail(c, BEM_ERR_HIP, std::string("mueller_grad: restoring B and X: ") +
                                                    hipGetErrorString(e));
            }
            return 0;
        }
        ~Restore() { (void)run(); }
    } restore{c, Bsr, Bsi, Xsr, Xsi, cnt};

    // 1. g = dL/dX into B (far-field adjoint), explicit dL/dka
    double expl_k = 0.0;
    if (int rc = ff_adjoint_launch(c, dL_dM, c->Br, c->Bi, &expl_k)) return rc;

    // 2. lambda = S Z^{-1} S g
    const long lo = (long)N * ldr, hi = (long)(2 * N) * ldr;
    const dim3 gneg((unsigned)((hi - lo + 255) / 256));
    hipLaunchKernelGGL(k_negate_lower, gneg, dim3(256), 0, s, lo, hi, c->Br, c->Bi);
    int iters = 0;
    double r0 = 0, r1 = 0, ts = 0, ti = 0;
    // refined to the tolerance the forward solution was refined to (1e-12
    // when that is not known): the gradient is no more accurate than X
    double lam_tol = c->x_tol > 1e-12 ? c->x_tol : 1e-12;
    if (const char* e = std::getenv("BEM_ADJOINT_TOL")) lam_tol = std::atof(e);
    const int solve_rc = lu_solve_launch(c, c->nrhs, lam_tol, 30, &iters, &r0, &r1, &ts, &ti);
    if (solve_rc < 0 && solve_rc != BEM_ERR_NOCONV) return solve_rc;
    BEM_HIP(c, hipMemcpyAsync(Lr, c->Xr, cnt * 8, hipMemcpyDeviceToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(Li, c->Xi, cnt * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_negate_lower, gneg, dim3(256), 0, s, lo, hi, Lr, Li);
    // restore the forward state
    if (int e = restore.run()) return e;

    // 3. W = Lambda X^T (npad x npad, K = ldr, a multiple of 16)
    hipLaunchKernelGGL(k_transpose, dim3((ldr + 31) / 32, (npad + 31) / 32), dim3(256), 0, s, npad, ldr, Xsr, Xsi,
                       (long)ldr, XTr, XTi, (long)npad);
    cgemm_ctx<double>(c, npad, npad, ldr, Lr, Li, ldr, XTr, XTi, npad, Wr, Wi, npad, 1);

    // 4. RHS term: sum lambda . dB/dka
    if (int e = rhs_dk_launch(c, dBr, dBi)) return e;
    hipLaunchKernelGGL(k_dot_rows, dim3(nblk_dot), dim3(256), 0, s, 2 * N, ldr, Lr, Li, dBr, dBi, (long)ldr, part);
    std::vector<double> hp(2 * (size_t)nblk_dot);
    BEM_HIP(c, hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s));
    BEM_HIP(c, hipStreamSynchronize(s));
    Cx rhs_term{0.0, 0.0};
    for (int b = 0; b < nblk_dot; b++) rhs_term = cadd(rhs_term, Cx{hp[2 * b], hp[2 * b + 1]});

    // 5. <dZ/dp, W> for p = Re m, Im m, ka
    const Cx k[2] = {{ka, 0.0}, {ap.kr[1], ap.ki[1]}};
    const double eta[2] = {1.0, ap.eta[1]};
    const double am3 = am * am * am;
    Cx kappa[3][2] = {{{0, 0}, {ka, 0}}, {{0, 0}, {0, ka}}, {{1, 0}, m}};
    double deta[3][2] = {{0, -m.r / am3}, {0, -m.i / am3}, {0, 0}};
    double dieta[3][2] = {{0, m.r / am}, {0, m.i / am}, {0, 0}};
    // One pass per moment kind forms the records of all three parameters
    // (two extra record arrays); without room for them, one pass per
    // parameter and kind.
    DevBuf recs;
    const size_t rec_bytes = asm_record_bytes(c);
    const bool merged = hipMalloc(&recs.p, 2 * rec_bytes) == hipSuccess;
    if (!merged) {
        recs.p = nullptr;
        (void)hipGetLastError();
    }
    Cx contr[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int pass = 0; pass < 2; pass++) {
        RecCoef q[3];
        for (int p = 0; p < 3; p++) q[p] = make_coef(pass, k, eta, kappa[p], deta[p], dieta[p]);
        double out[6];
        if (merged) {
            if (int e = asm_contract_launch(c, ap, q, 3, static_cast<double*>(recs.p), Wr, Wi, npad, out)) return e;
        } else {
            for (int p = 0; p < 3; p++)
                if (int e = asm_contract_launch(c, ap, q + p, 1, nullptr, Wr, Wi, npad, out + 2 * p)) return e;
        }
        for (int p = 0; p < 3; p++) contr[p] = cadd(contr[p], Cx{out[2 * p], out[2 * p + 1]});
    }
    for (int p = 0; p < 3; p++) grad[p] = -2.0 * contr[p].r;
    grad[2] += 2.0 * rhs_term.r + expl_k;
    BEM_HIP(c, hipGetLastError());
    if (solve_rc == BEM_ERR_NOCONV)  // the gradient is written, but lambda missed its tolerance
        return bem_fail(c, BEM_ERR_NOCONV,
                        "mueller_grad: the adjoint solve reached a relative residual of " + std::to_string(r1) +
                            ", above its tolerance " + std::to_string(lam_tol) + " (gradient written)");
    return 0;
}
