// farfield.hip -- plane-wave right-hand sides, far field, amplitude and
// Mueller matrices, orientation average.  FP64 throughout.
//
// Reference: compute_rhs_planewave (rhs.cpp:6-64), FFCache + farfield_batch_kernel
// + compute_farfield_batch_cuda (farfield.cu:9-346), the S-projection,
// amplitude_to_mueller and weighted average of main.cpp:809-846 and
// farfield.cu:464-508.
//
// Design.  The reference sums over the 2 N Nq half-points of every RWG half,
// one block per (call, direction), re-reading all quadrature data per block.
// Here the three halves on a triangle share its quadrature points, so the
// solution is first folded into a current density per triangle point
// (k_current: J_x = jw_x sum_a c_a (x - p_a) coef[n(a)], likewise M_x; 3x fewer
// points), and the phase e^{-ik r.x} of a point is shared by both polarisations
// of an orientation.  k_ff_partial stages a chunk of points in LDS and lets
// every thread of the block (one direction each) sweep it: the point data is
// read once per orientation, not once per direction.  Partial sums over point
// chunks are combined in a fixed order by k_ff_final (deterministic).
#include "../bem_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kChunk = 1024;  // points per k_ff_partial block
constexpr int kTile = 64;     // points per LDS tile

// ---------------------------------------------------------------------------
// RHS: B[n][j] = sum_halves sum_q f.E0_j e^{ik k_j.r} jw, B[N+n][j] likewise with
// H0_j = k_j x E0_j / eta.  prm: per RHS (k_hat, E0, H0) = 9 doubles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rhs(int N, int Nq, int npad, int ldr, int nrhs,
                                             const double4* __restrict__ qp, const double* __restrict__ hc,
                                             const double* __restrict__ hp, const int4* __restrict__ rwg_half,
                                             const double* __restrict__ prm, double kr, double ki,
                                             double* __restrict__ Br, double* __restrict__ Bi) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)npad * ldr) return;
    const int row = (int)(idx / ldr), j = (int)(idx % ldr);
    if (row >= N) {
        if (row >= 2 * N || j >= nrhs) {
            Br[idx] = 0.0;
            Bi[idx] = 0.0;
        }
        return;  // rows N..2N-1 are written by the thread of row - N
    }
    const long o2 = (long)(row + N) * ldr + j;
    if (j >= nrhs) {
        Br[idx] = Bi[idx] = Br[o2] = Bi[o2] = 0.0;
        return;
    }
    const double* P = prm + 9 * j;
    const double kx = P[0], ky = P[1], kz = P[2];
    const int4 h = rwg_half[row];
    double eR = 0, eI = 0, hR = 0, hI = 0;
    for (int half = 0; half < 2; half++) {
        const int t = half == 0 ? h.x : h.z, a = half == 0 ? h.y : h.w;
        const double ca = hc[t * 3 + a];
        const double* pa = hp + (t * 3 + a) * 3;
        double sEr = 0, sEi = 0, sHr = 0, sHi = 0;
        for (int q = 0; q < Nq; q++) {
            const double4 x = qp[t * Nq + q];
            const double fx = (x.x - pa[0]) * ca, fy = (x.y - pa[1]) * ca, fz = (x.z - pa[2]) * ca;
            const double sd = kx * x.x + ky * x.y + kz * x.z;
            double sn, cs;
            sincos(kr * sd, &sn, &cs);
            const double ea = exp(-ki * sd);
            const double pr = ea * cs * x.w, pi = ea * sn * x.w;
            const double fe = fx * P[3] + fy * P[4] + fz * P[5];
            const double fh = fx * P[6] + fy * P[7] + fz * P[8];
            sEr += fe * pr; sEi += fe * pi;
            sHr += fh * pr; sHi += fh * pi;
        }
        eR += sEr; eI += sEi;
        hR += sHr; hI += sHi;
    }
    Br[idx] = eR; Bi[idx] = eI;
    Br[o2] = hR;  Bi[o2] = hI;
}

// ---------------------------------------------------------------------------
// Current density per (call c, triangle point x): 12 doubles
// (Jx, Jy, Jz, Mx, My, Mz as (re, im)).  Coefficient of RWG n for call c at
// base[c * sc + n * sn] (J planes) and base[c * sc + (N + n) * sn] for M when
// `stacked` (device solutions), else separate M planes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_current(int T, int Nq, int n_calls, const double4* __restrict__ qp,
                                                 const double* __restrict__ hc, const double* __restrict__ hp,
                                                 const int* __restrict__ hrwg, const double* __restrict__ Jr,
                                                 const double* __restrict__ Ji, const double* __restrict__ Mr,
                                                 const double* __restrict__ Mi, long sc, long sn,
                                                 double* __restrict__ cur) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long P = (long)T * Nq;
    if (idx >= P * n_calls) return;
    const int c = (int)(idx / P);
    const long x = idx % P;
    const int t = (int)(x / Nq);
    const double4 r = qp[x];
    double v[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 3; a++) {
        const int n = hrwg[t * 3 + a];
        if (n < 0) continue;
        const double ca = hc[t * 3 + a] * r.w;
        const double* pa = hp + (t * 3 + a) * 3;
        const double f[3] = {(r.x - pa[0]) * ca, (r.y - pa[1]) * ca, (r.z - pa[2]) * ca};
        const long o = (long)c * sc + (long)n * sn;
        const double jr = Jr[o], ji = Ji[o], mr = Mr[o], mi = Mi[o];
        for (int e = 0; e < 3; e++) {
            v[2 * e] += f[e] * jr;
            v[2 * e + 1] += f[e] * ji;
            v[6 + 2 * e] += f[e] * mr;
            v[6 + 2 * e + 1] += f[e] * mi;
        }
    }
    double* out = cur + (c * P + x) * 12;
    for (int e = 0; e < 12; e++) out[e] = v[e];
}

// ---------------------------------------------------------------------------
// Partial far-field sums.  Block (o, chunk): thread d accumulates, over the
// chunk's points, e^{-ik rhat_{o,d}.x} times the currents of calls 2o, 2o+1.
// part[((o * nchunk + chunk) * ndir + d) * 24 + ...].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ff_partial(long P, int n_calls, int ndir, const double4* __restrict__ qp,
                                                    const double* __restrict__ cur, const double* __restrict__ rh,
                                                    double kr, double ki, double* __restrict__ part) {
    __shared__ double4 sx[kTile];
    __shared__ double sc[kTile][24];
    const int o = blockIdx.x, chunk = blockIdx.y, nchunk = gridDim.y;
    const long x0 = (long)chunk * kChunk, x1 = min(P, x0 + kChunk);
    const int c0 = 2 * o, ncall = min(2, n_calls - c0);
    for (int d0 = 0; d0 < ndir; d0 += 256) {
        const int d = d0 + threadIdx.x;
        double rx = 0, ry = 0, rz = 0;
        if (d < ndir) {
            const double* r = rh + ((long)o * ndir + d) * 3;
            rx = r[0]; ry = r[1]; rz = r[2];
        }
        double acc[24];
        for (int e = 0; e < 24; e++) acc[e] = 0.0;
        for (long xt = x0; xt < x1; xt += kTile) {
            const int nt = (int)min((long)kTile, x1 - xt);
            __syncthreads();
            for (int e = threadIdx.x; e < nt * 24; e += 256) {
                const int p = e / 24, k = e % 24;
                const int call = k / 12;
                sc[p][k] = call < ncall ? cur[((long)(c0 + call) * P + xt + p) * 12 + (k % 12)] : 0.0;
            }
            for (int p = threadIdx.x; p < nt; p += 256) sx[p] = qp[xt + p];
            __syncthreads();
            if (d < ndir) {
                for (int p = 0; p < nt; p++) {
                    const double rd = rx * sx[p].x + ry * sx[p].y + rz * sx[p].z;
                    double sn, cs;
                    sincos(kr * rd, &sn, &cs);
                    double pr = cs, pi = -sn;  // e^{-ik rd}
                    if (ki != 0.0) {
                        const double ea = exp(ki * rd);
                        pr *= ea;
                        pi *= ea;
                    }
#pragma unroll
                    for (int k = 0; k < 12; k++) {
                        const double ar = sc[p][2 * k], ai = sc[p][2 * k + 1];
                        acc[2 * k] = fma(pr, ar, fma(-pi, ai, acc[2 * k]));
                        acc[2 * k + 1] = fma(pr, ai, fma(pi, ar, acc[2 * k + 1]));
                    }
                }
            }
        }
        if (d < ndir) {
            double* out = part + (((long)o * nchunk + chunk) * ndir + d) * 24;
            for (int e = 0; e < 24; e++) out[e] = acc[e];
        }
    }
}

// F = -ik/(4 pi) [ eta (J - rhat (rhat.J)) - rhat x M ]   (farfield.cu:297-333)
__device__ __forceinline__ void far_vector(const double* j, const double* m, const double r[3], double kr, double ki,
                                           double eta, double F[6]) {
    const double rjr = r[0] * j[0] + r[1] * j[2] + r[2] * j[4];
    const double rji = r[0] * j[1] + r[1] * j[3] + r[2] * j[5];
    const double pfr = ki / (4.0 * kPi), pfi = -kr / (4.0 * kPi);
    for (int e = 0; e < 3; e++) {
        const int e1 = (e + 1) % 3, e2 = (e + 2) % 3;
        const double vr = eta * (j[2 * e] - r[e] * rjr) - (r[e1] * m[2 * e2] - r[e2] * m[2 * e1]);
        const double vi = eta * (j[2 * e + 1] - r[e] * rji) - (r[e1] * m[2 * e2 + 1] - r[e2] * m[2 * e1 + 1]);
        F[2 * e] = pfr * vr - pfi * vi;
        F[2 * e + 1] = pfr * vi + pfi * vr;
    }
}

// Sums the partials of (o, d) in chunk order and writes Fv[c][d][3] complex
// for the calls of orientation o.
__global__ __launch_bounds__(256) void k_ff_final(int n_orient, int n_calls, int ndir, int nchunk,
                                                  const double* __restrict__ part, const double* __restrict__ rh,
                                                  double kr, double ki, double eta, double* __restrict__ Fv) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_orient * ndir) return;
    const int o = idx / ndir, d = idx % ndir;
    double s[24];
    for (int e = 0; e < 24; e++) s[e] = 0.0;
    for (int ch = 0; ch < nchunk; ch++) {
        const double* p = part + (((long)o * nchunk + ch) * ndir + d) * 24;
        for (int e = 0; e < 24; e++) s[e] += p[e];
    }
    const double* r0 = rh + ((long)o * ndir + d) * 3;
    const double r[3] = {r0[0], r0[1], r0[2]};
    for (int call = 0; call < 2; call++) {
        const int c = 2 * o + call;
        if (c >= n_calls) break;
        double F[6];
        far_vector(s + 12 * call, s + 12 * call + 6, r, kr, ki, eta, F);
        double* out = Fv + ((long)c * ndir + d) * 6;
        for (int e = 0; e < 6; e++) out[e] = F[e];
    }
}

// Amplitudes and Mueller matrix per (orientation, angle), main.cpp:809-842 and
// farfield.cu:464-508; Mo[o][16][t] = M / |k|^2 (unweighted).
// geo per (o, t): rhat (3), e_par (3), e_perp (3).
__global__ __launch_bounds__(256) void k_mueller(int n_orient, int ntheta, const double* __restrict__ Fv,
                                                 const double* __restrict__ geo, double kr, double ki, int mirror,
                                                 double* __restrict__ Mo) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_orient * ntheta) return;
    const int o = idx / ntheta, t = idx % ntheta;
    const double* g = geo + (long)idx * 9;
    const double* Fp = Fv + ((long)(2 * o) * ntheta + t) * 6;
    const double* Fq = Fv + ((long)(2 * o + 1) * ntheta + t) * 6;
    auto proj = [](const double* F, const double* e, double& re, double& im) {
        re = F[0] * e[0] + F[2] * e[1] + F[4] * e[2];
        im = F[1] * e[0] + F[3] * e[1] + F[5] * e[2];
    };
    double a[4][2];  // F_par.ep, F_par.epp, F_perp.ep, F_perp.epp
    proj(Fp, g + 3, a[0][0], a[0][1]);
    proj(Fp, g + 6, a[1][0], a[1][1]);
    proj(Fq, g + 3, a[2][0], a[2][1]);
    proj(Fq, g + 6, a[3][0], a[3][1]);
    // S = -ik F with -ik = (ki, -kr)
    double S[5][2];  // S[1..4]
    const int map[4] = {2, 4, 3, 1};
    for (int u = 0; u < 4; u++) {
        S[map[u]][0] = ki * a[u][0] + kr * a[u][1];
        S[map[u]][1] = ki * a[u][1] - kr * a[u][0];
    }
    auto nrm = [&](int i) { return S[i][0] * S[i][0] + S[i][1] * S[i][1]; };
    auto mulc = [&](int i, int j, double& re, double& im) {  // S_i conj(S_j)
        re = S[i][0] * S[j][0] + S[i][1] * S[j][1];
        im = S[i][1] * S[j][0] - S[i][0] * S[j][1];
    };
    const double a1 = nrm(1), a2 = nrm(2), a3 = nrm(3), a4 = nrm(4);
    double s23r, s23i, s14r, s14i, s24r, s24i, s13r, s13i, s12r, s12i, s34r, s34i, s21r, s21i, s42r, s42i, s43r,
        s43i;
    mulc(2, 3, s23r, s23i);
    mulc(1, 4, s14r, s14i);
    mulc(2, 4, s24r, s24i);
    mulc(1, 3, s13r, s13i);
    mulc(1, 2, s12r, s12i);
    mulc(3, 4, s34r, s34i);
    mulc(2, 1, s21r, s21i);
    mulc(4, 2, s42r, s42i);
    mulc(4, 3, s43r, s43i);
    double M[16];
    M[0] = 0.5 * (a1 + a2 + a3 + a4);
    M[1] = 0.5 * (a2 - a1 + a4 - a3);
    M[2] = s23r + s14r;
    M[3] = s23i - s14i;
    M[4] = 0.5 * (a2 - a1 - a4 + a3);
    M[5] = 0.5 * (a2 + a1 - a4 - a3);
    M[6] = s23r - s14r;
    M[7] = s23i + s14i;
    M[8] = s24r + s13r;
    M[9] = s24r - s13r;
    M[10] = s12r + s34r;
    M[11] = s21i + s43i;
    M[12] = s42i + s13i;
    M[13] = s42i - s13i;
    M[14] = s12i - s34i;
    M[15] = s12r - s34r;
    if (mirror) {
        const int odd[8] = {2, 3, 6, 7, 8, 9, 12, 13};
        for (int b = 0; b < 8; b++) M[odd[b]] = 0.0;
    }
    const double k2 = kr * kr + ki * ki;
    for (int e = 0; e < 16; e++) Mo[((long)o * 16 + e) * ntheta + t] = M[e] / k2;
}

__global__ __launch_bounds__(256) void k_mueller_avg(int n_orient, int ntheta, const double* __restrict__ Mo,
                                                     const double* __restrict__ w, double* __restrict__ M) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 16 * ntheta) return;
    double s = 0.0;
    for (int o = 0; o < n_orient; o++) s += w[o] * Mo[(long)o * 16 * ntheta + idx];
    M[idx] = s;
}

// Far field of n_calls coefficient sets at the given directions into dFv.
int farfield_core(bem_ctx* c, const double* Jr, const double* Ji, const double* Mr, const double* Mi, long sc, long sn,
                  const double* drh, double kr, double ki, double eta, int n_calls, int n_orient, int ndir,
                  double* dFv, char* scratch) {
    const DevMesh& d = c->dm;
    const long P = (long)d.T * d.Nq;
    const int nchunk = (int)((P + kChunk - 1) / kChunk);
    double* cur = reinterpret_cast<double*>(scratch);
    double* part = cur + P * n_calls * 12;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_current, dim3((unsigned)((P * n_calls + 255) / 256)), dim3(256), 0, s, d.T, d.Nq, n_calls,
                       d.qp, d.hc, d.hp, d.hrwg, Jr, Ji, Mr, Mi, sc, sn, cur);
    const int kt = kt_mark(c, KS_FARFIELD);
    hipLaunchKernelGGL(k_ff_partial, dim3(n_orient, nchunk), dim3(256), 0, s, P, n_calls, ndir, d.qp, cur, drh, kr,
                       ki, part);
    kt_close(c, kt, (double)n_orient * ndir * P * (8.0 * 12.0), (double)n_orient * P * (4.0 + 24.0) * 8.0);
    hipLaunchKernelGGL(k_ff_final, dim3((n_orient * ndir + 255) / 256), dim3(256), 0, s, n_orient, n_calls, ndir,
                       nchunk, part, drh, kr, ki, eta, dFv);
    return 0;
}

size_t farfield_scratch_bytes(const DevMesh& d, int n_calls, int n_orient, int ndir) {
    const long P = (long)d.T * d.Nq;
    const int nchunk = (int)((P + kChunk - 1) / kChunk);
    return (size_t)(P * n_calls * 12 + (long)n_orient * nchunk * ndir * 24) * sizeof(double);
}

}  // namespace

int rhs_launch(bem_ctx* c, double kr, double ki, double eta, const double* khat, const double* E0, int nrhs) {
    std::vector<double> prm(9 * (size_t)nrhs);
    for (int j = 0; j < nrhs; j++) {
        const double* k = khat + 3 * j;
        const double* e = E0 + 3 * j;
        double* p = prm.data() + 9 * j;
        p[0] = k[0]; p[1] = k[1]; p[2] = k[2];
        p[3] = e[0]; p[4] = e[1]; p[5] = e[2];
        // H0 = k x E0 / eta (rhs.cpp:17)
        p[6] = (k[1] * e[2] - k[2] * e[1]) * (1.0 / eta);
        p[7] = (k[2] * e[0] - k[0] * e[2]) * (1.0 / eta);
        p[8] = (k[0] * e[1] - k[1] * e[0]) * (1.0 / eta);
    }
    void* s = nullptr;
    if (int rc = ctx_scratch(c, prm.size() * sizeof(double), &s)) return rc;
    BEM_HIP(c, hipMemcpyAsync(s, prm.data(), prm.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const DevMesh& d = c->dm;
    const long cnt = (long)c->npad * c->ldr;
    const int kt = kt_mark(c, KS_RHS);
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, d.N, d.Nq, c->npad,
                       c->ldr, nrhs, d.qp, d.hc, d.hp, d.rwg_half, static_cast<const double*>(s), kr, ki, c->Br,
                       c->Bi);
    kt_close(c, kt, (double)d.N * nrhs * 2 * d.Nq, (double)cnt * 16.0);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int ff_launch(bem_ctx* c, const double* dJr, const double* dJi, const double* dMr, const double* dMi, int ldc,
              const double* drh, double kr, double ki, double eta, int n_calls, int n_orient, int ndir,
              double* dFv) {
    // uploaded coefficients live in the scratch area already; allocate the
    // far-field work area separately
    const size_t bytes = farfield_scratch_bytes(c->dm, n_calls, n_orient, ndir);
    char* w = nullptr;
    if (hipMalloc(&w, bytes) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "farfield: work area");
    farfield_core(c, dJr, dJi, dMr, dMi, (long)ldc, 1, drh, kr, ki, eta, n_calls, n_orient, ndir, dFv, w);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(w);
    if (e != hipSuccess) return bem_fail(c, BEM_ERR_HIP, std::string("farfield: ") + hipGetErrorString(e));
    return 0;
}

int mueller_launch(bem_ctx* c, double kr, double ki, double eta, const double* RT, const double* w, int n_orient,
                   const double* theta, int ntheta, int mirror, double* M_avg, double* M_orient) {
    // directions and polarisation vectors in the body frame (main.cpp:766-799)
    std::vector<double> rh((size_t)n_orient * ntheta * 3), geo((size_t)n_orient * ntheta * 9);
    for (int o = 0; o < n_orient; o++) {
        const double* R = RT + 9 * o;
        auto mul = [R](const double v[3], double* out) {
            for (int i = 0; i < 3; i++) out[i] = R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
        };
        const double ey[3] = {0.0, 1.0, 0.0};
        double epp[3];
        mul(ey, epp);
        for (int t = 0; t < ntheta; t++) {
            const double ct = std::cos(theta[t]), st = std::sin(theta[t]);
            const double rl[3] = {st, 0.0, ct}, el[3] = {ct, 0.0, -st};
            double* g = geo.data() + ((size_t)o * ntheta + t) * 9;
            mul(rl, g);
            mul(el, g + 3);
            g[6] = epp[0]; g[7] = epp[1]; g[8] = epp[2];
            rh[((size_t)o * ntheta + t) * 3] = g[0];
            rh[((size_t)o * ntheta + t) * 3 + 1] = g[1];
            rh[((size_t)o * ntheta + t) * 3 + 2] = g[2];
        }
    }
    const int n_calls = 2 * n_orient;
    const size_t fvb = (size_t)n_calls * ntheta * 6 * sizeof(double);
    const size_t mob = (size_t)n_orient * 16 * ntheta * sizeof(double);
    const size_t small = rh.size() * 8 + geo.size() * 8 + (size_t)n_orient * 8 + 16 * (size_t)ntheta * 8;
    const size_t bytes = farfield_scratch_bytes(c->dm, n_calls, n_orient, ntheta) + fvb + mob + small + 1024;
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, bytes, &scr)) return rc;
    char* p = static_cast<char*>(scr);
    double* dFv = reinterpret_cast<double*>(p);
    double* dMo = dFv + (size_t)n_calls * ntheta * 6;
    double* drh = dMo + (size_t)n_orient * 16 * ntheta;
    double* dgeo = drh + rh.size();
    double* dw = dgeo + geo.size();
    double* dM = dw + n_orient;
    char* work = reinterpret_cast<char*>(dM + 16 * (size_t)ntheta);
    hipStream_t s = c->stream;
    BEM_HIP(c, hipMemcpyAsync(drh, rh.data(), rh.size() * 8, hipMemcpyHostToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(dgeo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(dw, w, (size_t)n_orient * 8, hipMemcpyHostToDevice, s));
    // device solution: column c is call c; J rows 0..N-1, M rows N..2N-1
    const int N = c->rwg.N;
    const long ld = c->ldr;
    farfield_core(c, c->Xr, c->Xi, c->Xr + (long)N * ld, c->Xi + (long)N * ld, 1, ld, drh, kr, ki, eta, n_calls,
                  n_orient, ntheta, dFv, work);
    hipLaunchKernelGGL(k_mueller, dim3((n_orient * ntheta + 255) / 256), dim3(256), 0, s, n_orient, ntheta, dFv, dgeo,
                       kr, ki, mirror, dMo);
    hipLaunchKernelGGL(k_mueller_avg, dim3((16 * ntheta + 255) / 256), dim3(256), 0, s, n_orient, ntheta, dMo, dw, dM);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipMemcpyAsync(M_avg, dM, 16 * (size_t)ntheta * 8, hipMemcpyDeviceToHost, s));
    if (M_orient) BEM_HIP(c, hipMemcpyAsync(M_orient, dMo, mob, hipMemcpyDeviceToHost, s));
    BEM_HIP(c, hipStreamSynchronize(s));
    return 0;
}
