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
// dk = 1: the derivative with respect to k (phase factor i k_hat.r).
__global__ __launch_bounds__(256) void k_rhs(int N, int Nq, int npad, int ldr, int nrhs,
                                             const double4* __restrict__ qp, const double* __restrict__ hc,
                                             const double* __restrict__ hp, const int4* __restrict__ rwg_half,
                                             const double* __restrict__ prm, double kr, double ki, int dk,
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
            double pr = ea * cs * x.w, pi = ea * sn * x.w;
            if (dk) {
                const double t0 = -sd * pi;
                pi = sd * pr;
                pr = t0;
            }
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
                  double* dFv, char* work) {
    const DevMesh& d = c->dm;
    const long P = (long)d.T * d.Nq;
    const int nchunk = (int)((P + kChunk - 1) / kChunk);
    double* cur = reinterpret_cast<double*>(work);
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

// ---------------------------------------------------------------------------
// Adjoint of the far-field chain (holomorphic derivatives: for the real loss
// L, dL = 2 Re sum (dL/dv) dv).  Per (orientation o, angle t), given
// G = dL/dM_avg: sigma_l = dL/dS_l from the Mueller quadratic forms with
// weights G_e w_o / |k|^2, then dL/dF (S = -ik F.e), dL/dV (F = -ik/(4 pi) V)
// and coef = (dL/dJ~, dL/dM~) of both polarisations with
// V = eta (J~ - rhat (rhat.J~)) - rhat x M~.  expl[o][t] receives the explicit
// dL/dk of this (o, t) through the prefactors: 2 Re sum sigma_l (2/k) S_l
// (S = -k^2/(4 pi) V.e) and -(2/k) sum G_e w_o M_e / |k|^2 (the 1/|k|^2
// normalisation; k = k_ext real).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_mueller_bwd(int n_orient, int ntheta, const double* __restrict__ Fv,
                                                     const double* __restrict__ geo, const double* __restrict__ G,
                                                     const double* __restrict__ w, double kr, double ki, double eta,
                                                     int mirror, double* __restrict__ coef,
                                                     double* __restrict__ expl) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_orient * ntheta) return;
    const int o = idx / ntheta, t = idx % ntheta;
    const double* g = geo + (long)idx * 9;
    const double* F2[2] = {Fv + ((long)(2 * o) * ntheta + t) * 6, Fv + ((long)(2 * o + 1) * ntheta + t) * 6};
    auto proj = [](const double* F, const double* e, double* out) {
        out[0] = F[0] * e[0] + F[2] * e[1] + F[4] * e[2];
        out[1] = F[1] * e[0] + F[3] * e[1] + F[5] * e[2];
    };
    double a[4][2];  // F_par.ep, F_par.epp, F_perp.ep, F_perp.epp
    proj(F2[0], g + 3, a[0]);
    proj(F2[0], g + 6, a[1]);
    proj(F2[1], g + 3, a[2]);
    proj(F2[1], g + 6, a[3]);
    const double mik[2] = {ki, -kr};  // -ik
    double S[5][2];
    const int map[4] = {2, 4, 3, 1};
    for (int u = 0; u < 4; u++) {
        S[map[u]][0] = mik[0] * a[u][0] - mik[1] * a[u][1];
        S[map[u]][1] = mik[0] * a[u][1] + mik[1] * a[u][0];
    }
    const double k2 = kr * kr + ki * ki;
    double cw[16];
    for (int e = 0; e < 16; e++) cw[e] = G[(long)e * ntheta + t] * w[o] / k2;
    if (mirror) {
        const int odd[8] = {2, 3, 6, 7, 8, 9, 12, 13};
        for (int b = 0; b < 8; b++) cw[odd[b]] = 0.0;
    }
    // coefficients of |S_i|^2, Re(S_i conj S_j), Im(S_i conj S_j) in sum_e cw_e M_e
    const double A1 = 0.5 * (cw[0] - cw[1] - cw[4] + cw[5]), A2 = 0.5 * (cw[0] + cw[1] + cw[4] + cw[5]);
    const double A3 = 0.5 * (cw[0] - cw[1] + cw[4] - cw[5]), A4 = 0.5 * (cw[0] + cw[1] - cw[4] - cw[5]);
    double sig[5][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    auto quad_abs = [&](int i, double cf) {  // cf |S_i|^2
        sig[i][0] += cf * S[i][0];
        sig[i][1] -= cf * S[i][1];
    };
    auto quad_re = [&](int i, int j, double cf) {  // cf Re(S_i conj S_j)
        sig[i][0] += 0.5 * cf * S[j][0];
        sig[i][1] -= 0.5 * cf * S[j][1];
        sig[j][0] += 0.5 * cf * S[i][0];
        sig[j][1] -= 0.5 * cf * S[i][1];
    };
    auto quad_im = [&](int i, int j, double cf) {  // cf Im(S_i conj S_j): d/dS_i = conj S_j/(2i), d/dS_j = -conj S_i/(2i)
        // conj(S)/(2i) = (-i/2)(Sr - i Si) = (-Si/2, -Sr/2)
        sig[i][0] += cf * (-0.5 * S[j][1]);
        sig[i][1] += cf * (-0.5 * S[j][0]);
        sig[j][0] -= cf * (-0.5 * S[i][1]);
        sig[j][1] -= cf * (-0.5 * S[i][0]);
    };
    quad_abs(1, A1);
    quad_abs(2, A2);
    quad_abs(3, A3);
    quad_abs(4, A4);
    quad_re(2, 3, cw[2] + cw[6]);
    quad_re(1, 4, cw[2] - cw[6]);
    quad_im(2, 3, cw[3] + cw[7]);
    quad_im(1, 4, cw[7] - cw[3]);
    quad_re(2, 4, cw[8] + cw[9]);
    quad_re(1, 3, cw[8] - cw[9]);
    quad_re(1, 2, cw[10] + cw[15]);
    quad_re(3, 4, cw[10] - cw[15]);
    quad_im(2, 1, cw[11]);
    quad_im(4, 3, cw[11]);
    quad_im(4, 2, cw[12] + cw[13]);
    quad_im(1, 3, cw[12] - cw[13]);
    quad_im(1, 2, cw[14]);
    quad_im(3, 4, -cw[14]);
    // explicit dL/dk: 2 Re sum sigma (2/k) S  and  -(2/k) sum cw M
    double sS[2] = {0.0, 0.0}, cM = 0.0;
    for (int l = 1; l <= 4; l++) {
        sS[0] += sig[l][0] * S[l][0] - sig[l][1] * S[l][1];
        sS[1] += sig[l][0] * S[l][1] + sig[l][1] * S[l][0];
    }
    {
        const double a1 = S[1][0] * S[1][0] + S[1][1] * S[1][1], a2 = S[2][0] * S[2][0] + S[2][1] * S[2][1];
        const double a3 = S[3][0] * S[3][0] + S[3][1] * S[3][1], a4 = S[4][0] * S[4][0] + S[4][1] * S[4][1];
        cM = A1 * a1 + A2 * a2 + A3 * a3 + A4 * a4;
        auto re_ij = [&](int i, int j) { return S[i][0] * S[j][0] + S[i][1] * S[j][1]; };
        auto im_ij = [&](int i, int j) { return S[i][1] * S[j][0] - S[i][0] * S[j][1]; };
        cM += (cw[2] + cw[6]) * re_ij(2, 3) + (cw[2] - cw[6]) * re_ij(1, 4) + (cw[3] + cw[7]) * im_ij(2, 3) +
              (cw[7] - cw[3]) * im_ij(1, 4) + (cw[8] + cw[9]) * re_ij(2, 4) + (cw[8] - cw[9]) * re_ij(1, 3) +
              (cw[10] + cw[15]) * re_ij(1, 2) + (cw[10] - cw[15]) * re_ij(3, 4) + cw[11] * (im_ij(2, 1) + im_ij(4, 3)) +
              (cw[12] + cw[13]) * im_ij(4, 2) + (cw[12] - cw[13]) * im_ij(1, 3) + cw[14] * (im_ij(1, 2) - im_ij(3, 4));
    }
    // 2/k (complex)
    const double tk[2] = {2.0 * kr / k2, -2.0 * ki / k2};
    expl[idx] = 2.0 * (tk[0] * sS[0] - tk[1] * sS[1]) - tk[0] * cM;
    // dL/dF of both calls, then dL/dV = (-ik/(4 pi)) dL/dF
    const double* r = g;
    const double* ep = g + 3;
    const double* epp = g + 6;
    const int lp[2][2] = {{2, 4}, {3, 1}};  // call 0: S2 (e_par), S4 (e_perp); call 1: S3, S1
    const double pv[2] = {mik[0] / (4.0 * kPi), mik[1] / (4.0 * kPi)};
    for (int call = 0; call < 2; call++) {
        double dF[3][2];
        for (int e = 0; e < 3; e++) {
            const double* s1 = sig[lp[call][0]];
            const double* s2 = sig[lp[call][1]];
            const double vr = s1[0] * ep[e] + s2[0] * epp[e], vi = s1[1] * ep[e] + s2[1] * epp[e];
            dF[e][0] = mik[0] * vr - mik[1] * vi;
            dF[e][1] = mik[0] * vi + mik[1] * vr;
        }
        double dV[3][2];
        for (int e = 0; e < 3; e++) {
            dV[e][0] = pv[0] * dF[e][0] - pv[1] * dF[e][1];
            dV[e][1] = pv[0] * dF[e][1] + pv[1] * dF[e][0];
        }
        const double rvr = r[0] * dV[0][0] + r[1] * dV[1][0] + r[2] * dV[2][0];
        const double rvi = r[0] * dV[0][1] + r[1] * dV[1][1] + r[2] * dV[2][1];
        double* out = coef + (long)idx * 24 + call * 12;
        for (int e = 0; e < 3; e++) {
            const int e1 = (e + 1) % 3, e2 = (e + 2) % 3;
            out[2 * e] = eta * (dV[e][0] - r[e] * rvr);
            out[2 * e + 1] = eta * (dV[e][1] - r[e] * rvi);
            out[6 + 2 * e] = r[e1] * dV[e2][0] - r[e2] * dV[e1][0];  // rhat x dL/dV
            out[6 + 2 * e + 1] = r[e1] * dV[e2][1] - r[e2] * dV[e1][1];
        }
    }
}

// Transpose of the far-field sums.  Block (o, chunk of 256 points), thread =
// point x: A_c(x) = sum_d coef_{o,d,c} e^{-ik rhat_d.x} (dL/d current at x of
// call c = 2o, 2o+1) into acur, and the explicit phase term
// 2 Re sum_c cur_c(x) . sum_d coef_{o,d,c} (-i rhat_d.x) e^{-ik rhat_d.x}
// summed over the block into epart[o * gridDim.y + chunk].
constexpr int kDirTile = 64;
__global__ __launch_bounds__(256) void k_ff_adj(long P, int n_calls, int ndir, const double4* __restrict__ qp,
                                                const double* __restrict__ cur, const double* __restrict__ coef,
                                                const double* __restrict__ rh, double kr, double ki,
                                                double* __restrict__ acur, double* __restrict__ epart) {
    __shared__ double sc[kDirTile][24];
    __shared__ double sr[kDirTile][3];
    __shared__ double red[256];
    const int o = blockIdx.x;
    const long x = (long)blockIdx.y * 256 + threadIdx.x;
    const bool ok = x < P;
    double4 X = ok ? qp[x] : make_double4(0, 0, 0, 0);
    double A[24], Ad[24];
    for (int e = 0; e < 24; e++) A[e] = Ad[e] = 0.0;
    for (int d0 = 0; d0 < ndir; d0 += kDirTile) {
        const int nd = min(kDirTile, ndir - d0);
        __syncthreads();
        for (int e = threadIdx.x; e < nd * 24; e += 256)
            sc[e / 24][e % 24] = coef[((long)o * ndir + d0 + e / 24) * 24 + e % 24];
        for (int e = threadIdx.x; e < nd * 3; e += 256) sr[e / 3][e % 3] = rh[((long)o * ndir + d0 + e / 3) * 3 + e % 3];
        __syncthreads();
        if (ok) {
            for (int d = 0; d < nd; d++) {
                const double rd = sr[d][0] * X.x + sr[d][1] * X.y + sr[d][2] * X.z;
                double sn, cs;
                sincos(kr * rd, &sn, &cs);
                double pr = cs, pi = -sn;  // e^{-ik rd}
                if (ki != 0.0) {
                    const double ea = exp(ki * rd);
                    pr *= ea;
                    pi *= ea;
                }
                const double qr = rd * pi, qi = -rd * pr;  // (-i rd) e^{-ik rd}
#pragma unroll
                for (int k = 0; k < 12; k++) {
                    const double ar = sc[d][2 * k], ai = sc[d][2 * k + 1];
                    A[2 * k] = fma(pr, ar, fma(-pi, ai, A[2 * k]));
                    A[2 * k + 1] = fma(pr, ai, fma(pi, ar, A[2 * k + 1]));
                    Ad[2 * k] = fma(qr, ar, fma(-qi, ai, Ad[2 * k]));
                    Ad[2 * k + 1] = fma(qr, ai, fma(qi, ar, Ad[2 * k + 1]));
                }
            }
        }
    }
    double ex = 0.0;
    if (ok) {
        for (int call = 0; call < 2; call++) {
            const int c = 2 * o + call;
            if (c >= n_calls) break;
            const double* cu = cur + ((long)c * P + x) * 12;
            double* out = acur + ((long)c * P + x) * 12;
            for (int k = 0; k < 6; k++) {
                const double ar = A[call * 12 + 2 * k], ai = A[call * 12 + 2 * k + 1];
                out[2 * k] = ar;
                out[2 * k + 1] = ai;
                ex += cu[2 * k] * Ad[call * 12 + 2 * k] - cu[2 * k + 1] * Ad[call * 12 + 2 * k + 1];
            }
        }
    }
    red[threadIdx.x] = 2.0 * ex;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) epart[(long)o * gridDim.y + blockIdx.y] = red[0];
}

// Transpose of k_current: g[n][c] = sum over the halves (t, a) of RWG n and
// the points x of t of f_a(x) jw_x . A_{c,J}(x); g[N + n][c] likewise with A_M.
__global__ __launch_bounds__(256) void k_current_adj(int N, int Nq, int n_calls, int ldr,
                                                     const double4* __restrict__ qp, const double* __restrict__ hc,
                                                     const double* __restrict__ hp, const int4* __restrict__ rwg_half,
                                                     const double* __restrict__ acur, long P,
                                                     double* __restrict__ gr, double* __restrict__ gi) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)N * n_calls) return;
    const int n = (int)(idx / n_calls), c = (int)(idx % n_calls);
    const int4 h = rwg_half[n];
    double J[2] = {0, 0}, M[2] = {0, 0};
    for (int half = 0; half < 2; half++) {
        const int t = half == 0 ? h.x : h.z, a = half == 0 ? h.y : h.w;
        const double ca = hc[t * 3 + a];
        const double* pa = hp + (t * 3 + a) * 3;
        for (int q = 0; q < Nq; q++) {
            const long x = (long)t * Nq + q;
            const double4 r = qp[x];
            const double f[3] = {(r.x - pa[0]) * ca * r.w, (r.y - pa[1]) * ca * r.w, (r.z - pa[2]) * ca * r.w};
            const double* A = acur + ((long)c * P + x) * 12;
            for (int e = 0; e < 3; e++) {
                J[0] += f[e] * A[2 * e];
                J[1] += f[e] * A[2 * e + 1];
                M[0] += f[e] * A[6 + 2 * e];
                M[1] += f[e] * A[6 + 2 * e + 1];
            }
        }
    }
    gr[(long)n * ldr + c] = J[0];
    gi[(long)n * ldr + c] = J[1];
    gr[(long)(N + n) * ldr + c] = M[0];
    gi[(long)(N + n) * ldr + c] = M[1];
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
                       c->ldr, nrhs, d.qp, d.hc, d.hp, d.rwg_half, static_cast<const double*>(s), kr, ki, 0, c->Br,
                       c->Bi);
    kt_close(c, kt, (double)d.N * nrhs * 2 * d.Nq, (double)cnt * 16.0);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    c->rhs_prm.swap(prm);
    c->rhs_k[0] = kr;
    c->rhs_k[1] = ki;
    c->rhs_eta = eta;
    return 0;
}

int rhs_dk_launch(bem_ctx* c, double* dBr, double* dBi) {
    const int nrhs = (int)(c->rhs_prm.size() / 9);
    if (nrhs < 1 || nrhs > c->ldr) return bem_fail(c, BEM_ERR_STATE, "rhs derivative: no plane waves stored");
    double* s = nullptr;
    if (hipMalloc(&s, c->rhs_prm.size() * sizeof(double)) != hipSuccess)
        return bem_fail(c, BEM_ERR_NOMEM, "rhs derivative: parameters");
    BEM_HIP(c, hipMemcpy(s, c->rhs_prm.data(), c->rhs_prm.size() * sizeof(double), hipMemcpyHostToDevice));
    const DevMesh& d = c->dm;
    const long cnt = (long)c->npad * c->ldr;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, d.N, d.Nq, c->npad,
                       c->ldr, nrhs, d.qp, d.hc, d.hp, d.rwg_half, s, c->rhs_k[0], c->rhs_k[1], 1, dBr, dBi);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(s);
    if (e != hipSuccess) return bem_fail(c, BEM_ERR_HIP, std::string("rhs derivative: ") + hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------------------------
// The far field without device allocation (bem_farfield_noalloc): every
// device buffer of the call lives in `scratch`, one static array per GPU that
// the library holds until the process ends.  Host code never uses the array's
// name as an address (it is a host-side stand-in): its device address comes
// from hipGetSymbolAddress once per context, with the context's device
// current, and only that address is passed to kernels and copies.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[kNoAllocBytes];

int ff_noalloc_init(bem_ctx* c) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->noalloc_ptr = nullptr;
        return bem_fail(c, BEM_ERR_HIP, std::string("no-allocation scratch: ") + hipGetErrorString(e));
    }
    c->noalloc_ptr = p;
    return 0;
}

int ff_noalloc_launch(bem_ctx* c, const double* cJ, const double* cM, const double* rh, double kr, double ki,
                      double eta, int n_calls, int n_orient, int ndir, double* Fv) {
    if (!c->noalloc_ptr) return bem_fail(c, BEM_ERR_STATE, "farfield_noalloc: no device address of the scratch array");
    const int N = c->rwg.N;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bc = up((size_t)n_calls * N * 2 * sizeof(double));       // one interleaved coefficient set
    const size_t br = up((size_t)n_orient * ndir * 3 * sizeof(double));    // directions
    const size_t bf = up((size_t)n_calls * ndir * 6 * sizeof(double));     // result
    const size_t bw = up(farfield_scratch_bytes(c->dm, n_calls, n_orient, ndir));  // currents, partial sums
    const size_t need = 2 * bc + br + bf + bw;
    if (need > kNoAllocBytes)
        return bem_fail(c, BEM_ERR_ARG, "farfield_noalloc: the call needs " + std::to_string(need) +
                                            " bytes of device staging, above the " + std::to_string(kNoAllocBytes) +
                                            "-byte no-allocation array (use bem_farfield)");
    char* base = static_cast<char*>(c->noalloc_ptr);
    double* dJ = reinterpret_cast<double*>(base);
    double* dM = reinterpret_cast<double*>(base + bc);
    double* drh = reinterpret_cast<double*>(base + 2 * bc);
    double* dFv = reinterpret_cast<double*>(base + 2 * bc + br);
    char* work = base + 2 * bc + br + bf;
    hipStream_t s = c->stream;
    BEM_HIP(c, hipMemcpyAsync(dJ, cJ, (size_t)n_calls * N * 2 * sizeof(double), hipMemcpyHostToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(dM, cM, (size_t)n_calls * N * 2 * sizeof(double), hipMemcpyHostToDevice, s));
    BEM_HIP(c, hipMemcpyAsync(drh, rh, (size_t)n_orient * ndir * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    // the coefficients stay interleaved: call c, RWG n at 2 (c N + n) (+1 imaginary)
    farfield_core(c, dJ, dJ + 1, dM, dM + 1, 2L * N, 2L, drh, kr, ki, eta, n_calls, n_orient, ndir, dFv, work);
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipMemcpyAsync(Fv, dFv, (size_t)n_calls * ndir * 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    BEM_HIP(c, hipStreamSynchronize(s));
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

// Directions rhat (n_orient x ntheta x 3) and (rhat, e_par, e_perp) per
// (orientation, angle) in the body frame (main.cpp:766-799).
static void build_dirs(const double* RT, int n_orient, const double* theta, int ntheta, std::vector<double>& rh,
                       std::vector<double>& geo) {
    rh.assign((size_t)n_orient * ntheta * 3, 0.0);
    geo.assign((size_t)n_orient * ntheta * 9, 0.0);
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
}

int mueller_launch(bem_ctx* c, double kr, double ki, double eta, const double* RT, const double* w, int n_orient,
                   const double* theta, int ntheta, int mirror, double* M_avg, double* M_orient) {
    std::vector<double> rh, geo;
    build_dirs(RT, n_orient, theta, ntheta, rh, geo);
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

int ff_adjoint_launch(bem_ctx* c, const double* dL_dM, double* gr, double* gi, double* explicit_dk) {
    const int n_orient = c->last_norient, ntheta = c->last_ntheta, n_calls = 2 * n_orient;
    const double kr = c->last_kext[0], ki = c->last_kext[1], eta = c->last_eta_ext;
    std::vector<double> rh, geo;
    build_dirs(c->last_RT.data(), n_orient, c->last_theta.data(), ntheta, rh, geo);
    const DevMesh& d = c->dm;
    const long P = (long)d.T * d.Nq;
    const int nchunk = (int)((P + 255) / 256);
    const size_t nfv = (size_t)n_calls * ntheta * 6, ncoef = (size_t)n_orient * ntheta * 24;
    const size_t nacur = (size_t)n_calls * P * 12, nsmall = rh.size() + geo.size() + (size_t)n_orient +
                                                           16 * (size_t)ntheta + (size_t)n_orient * ntheta +
                                                           (size_t)n_orient * nchunk;
    const size_t bytes = farfield_scratch_bytes(d, n_calls, n_orient, ntheta) +
                         (nfv + ncoef + nacur + nsmall + 64) * sizeof(double);
    char* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "far-field adjoint: work area");
    double* work = reinterpret_cast<double*>(base);  // cur + partials of farfield_core first
    double* dFv = work + farfield_scratch_bytes(d, n_calls, n_orient, ntheta) / sizeof(double);
    double* dcoef = dFv + nfv;
    double* dacur = dcoef + ncoef;
    double* drh = dacur + nacur;
    double* dgeo = drh + rh.size();
    double* dw = dgeo + geo.size();
    double* dG = dw + n_orient;
    double* dexpl = dG + 16 * (size_t)ntheta;
    double* depart = dexpl + (size_t)n_orient * ntheta;
    hipStream_t s = c->stream;
    int rc = 0;
    auto fail = [&](hipError_t e, const char* what) {
        rc = bem_fail(c, BEM_ERR_HIP, std::string("far-field adjoint: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipMemcpyAsync(drh, rh.data(), rh.size() * 8, hipMemcpyHostToDevice, s);
    if (!e) e = hipMemcpyAsync(dgeo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice, s);
    if (!e) e = hipMemcpyAsync(dw, c->last_w.data(), (size_t)n_orient * 8, hipMemcpyHostToDevice, s);
    if (!e) e = hipMemcpyAsync(dG, dL_dM, 16 * (size_t)ntheta * 8, hipMemcpyHostToDevice, s);
    if (!e) e = hipMemsetAsync(gr, 0, (size_t)c->npad * c->ldr * 8, s);
    if (!e) e = hipMemsetAsync(gi, 0, (size_t)c->npad * c->ldr * 8, s);
    if (e) fail(e, "upload");
    std::vector<double> hexpl((size_t)n_orient * ntheta), hep((size_t)n_orient * nchunk);
    if (!rc) {
        const int N = c->rwg.N;
        const long ld = c->ldr;
        // forward far field again (currents in `work`, field in dFv)
        farfield_core(c, c->Xr, c->Xi, c->Xr + (long)N * ld, c->Xi + (long)N * ld, 1, ld, drh, kr, ki, eta, n_calls,
                      n_orient, ntheta, dFv, reinterpret_cast<char*>(work));
        hipLaunchKernelGGL(k_mueller_bwd, dim3((n_orient * ntheta + 255) / 256), dim3(256), 0, s, n_orient, ntheta,
                           dFv, dgeo, dG, dw, kr, ki, eta, c->last_mirror, dcoef, dexpl);
        hipLaunchKernelGGL(k_ff_adj, dim3(n_orient, nchunk), dim3(256), 0, s, P, n_calls, ntheta, d.qp, work, dcoef,
                           drh, kr, ki, dacur, depart);
        hipLaunchKernelGGL(k_current_adj, dim3((unsigned)(((long)N * n_calls + 255) / 256)), dim3(256), 0, s, N,
                           d.Nq, n_calls, c->ldr, d.qp, d.hc, d.hp, d.rwg_half, dacur, P, gr, gi);
        e = hipMemcpyAsync(hexpl.data(), dexpl, hexpl.size() * 8, hipMemcpyDeviceToHost, s);
        if (!e) e = hipMemcpyAsync(hep.data(), depart, hep.size() * 8, hipMemcpyDeviceToHost, s);
        if (!e) e = hipStreamSynchronize(s);
        if (e) fail(e, "kernels");
    }
    (void)hipFree(base);
    if (rc) return rc;
    double sum = 0.0;
    for (double v : hexpl) sum += v;
    for (double v : hep) sum += v;
    *explicit_dk = sum;
    return 0;
}
