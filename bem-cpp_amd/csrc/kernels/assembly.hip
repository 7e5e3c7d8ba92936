// assembly.hip -- PMCHWT matrix assembly on gfx950.
//
// Reference algorithm: assemble_LK_kernel + precompute_half + the three
// pp/pm/mm passes + apply_singular_corrections + symmetrisation + block
// formation (assembly.cu:14-135, 150-201, 208-321, 328-505; pmchwt.cu:6-56).
//
// Design.  The reference evaluates every (test RWG half, source RWG half) pair
// on its own: 3 N^2 half pairs x Nq^2 point pairs per wavenumber.  All halves
// on a triangle share its quadrature points, so the Green's function values of
// a triangle pair (T, T') serve all 3 x 3 half pairs on it.  With
//   f_a(r) = c_a (r - p_a)            (c_a = +-l/(2A), p_a free vertex)
// the half-pair integrals are fixed combinations of 14 complex moments of the
// triangle pair (W_ij = jw_i jw_j, G = e^{ikR}/(4 pi R), g = G (ik - 1/R)/R):
//   M0 = sum W G,  Mr = sum W G r_i,  Ms = sum W G r_j,  Mrs = sum W G r_i.r_j
//   Nx = sum W g r_i x r_j,           Nd = sum W g (r_i - r_j)
//   L_ab = c_a c_b [ ik (Mrs - p_a.Ms - p_b.Mr + (p_a.p_b) M0) - (4i/k) M0 ]
//   K_ab = c_a c_b [ (p_b - p_a).Nx + (p_b x p_a).Nd ]
// (the K identity: (r_i-p_a).((r_i-r_j) x (r_j-p_b)) = (p_b-p_a).(r_i x r_j)
//  + (p_b x p_a).(r_i - r_j)).  One pass over the T(T+1)/2 triangle pairs,
// Nq^2 point pairs each, both wavenumbers sharing the geometry, replaces the
// reference's 6 N^2 Nq^2 half-pair evaluations (13.5x fewer point pairs).
//
// Precision.  Distances and phases are formed in FP64 (FP32 rsqrt seed + one
// FP64 Newton step), the phase is reduced to a quarter turn in FP64, and only
// sin/cos/exp of the reduced argument run in FP32; all moments accumulate in
// FP64 (their expansion about the origin is exact for FP64 sums of the same
// FP32-rounded G values).  Triangle pairs that share a vertex (self and
// adjacent: singular and near-singular integrands) run entirely in FP64 in
// their own kernel, which also adds the Graglia static part for self pairs.
//
// Kernels, in launch order:
//   k_graglia      P(r_q), V(r_q) at every triangle's own points      (FP64)
//   k_tripair_far  non-touching triangle pairs -> H slots     (FP32 trig, FP64 sums)
//   k_tripair_near touching pairs incl. self    -> H slots     (FP64)
//   k_gather       Z[m][n] = Z[n][m] = sum of the 4 half-pair slots of (m, n)
//                  in a fixed order; PMCHWT blocks formed in place, FP64 and FP32
//   k_pad          identity on the padding rows/columns
// Half-pair value (a on t, b on s): {P, Q, Ks} complex (6 doubles),
// P = eta_e L_e + eta_i L_i, Q = L_e/eta_e + L_i/eta_i, Ks = K_e + K_i.
// A triangle pair's 9 values go to the RWG pairs that read them: H holds, for
// every RWG pair m <= n, the 4 slots of its half combinations (hm, hn) side
// by side (192 bytes), grouped by 16 x 16 tiles of (m, n) (slot_index), so
// the gather and the contraction read each slot once with coalesced loads and
// the scatter is paid by the VALU-bound producers.
// No atomics: every slot has one writer, every Z entry one reader-summer.
#include "../bem_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv4Pi = 1.0 / (4.0 * kPi);
constexpr int kRec = 54;  // doubles of the 9 half-pair values of one triangle pair
constexpr int kSlot = 6;  // doubles per half-pair slot: P, Q, Ks complex
constexpr int kGt = 16;   // gather / contraction tile edge (RWGs)

// Slot of RWG pair (m <= n), halves hm of m and hn of n (0: plus, 1: minus):
// the 4 slots of a pair together, the pairs of a 16 x 16 tile of (m, n) in
// row-major order, the tiles with m tile <= n tile packed by tile rows.
__device__ __forceinline__ long long slot_index(int m, int n, int hm, int hn, int ntile) {
    const long long bi = m / kGt, bj = n / kGt;
    const long long tile = bi * ntile - bi * (bi - 1) / 2 + (bj - bi);
    return (tile * (kGt * kGt) + (m % kGt) * kGt + (n % kGt)) * 4 + hm * 2 + hn;
}

// Slot that reads the value of (a on t, b on s), t <= s, or -1 when none does
// (a self pair's (a, b) with RWG m > n: its (b, a) value serves the pair).
// The two halves of one RWG on two triangles (m == n, t != s) give both half
// orders of (m, m) the same value: *dup receives the second slot.
__device__ __forceinline__ long long half_pair_slot(const int* __restrict__ hrwg, const double* __restrict__ hc,
                                                    int t, int a, int s, int b, int ntile, long long* dup) {
    *dup = -1;
    const int m = hrwg[t * 3 + a], n = hrwg[s * 3 + b];
    if (m < 0 || n < 0) return -1;
    const int hm = hc[t * 3 + a] < 0.0, hn = hc[s * 3 + b] < 0.0;  // c < 0 on the minus half
    if (m < n) return slot_index(m, n, hm, hn, ntile);
    if (m > n) return t == s ? -1 : slot_index(n, m, hn, hm, ntile);
    if (t != s) *dup = slot_index(m, m, hn, hm, ntile);
    return slot_index(m, m, hm, hn, ntile);
}

__device__ __forceinline__ void store_slot(double* __restrict__ H, long long slot, long long dup, double2 v0,
                                           double2 v1, double2 v2) {
    double2* o = reinterpret_cast<double2*>(H + slot * kSlot);
    o[0] = v0;
    o[1] = v1;
    o[2] = v2;
    if (dup >= 0) {
        o = reinterpret_cast<double2*>(H + dup * kSlot);
        o[0] = v0;
        o[1] = v1;
        o[2] = v2;
    }
}

// sin/cos on [-pi/4, pi/4]: Taylor series to x^9 / x^10 (truncation < 2e-9).
__device__ __forceinline__ void sincos_poly(float x, float& s, float& c) {
    const float x2 = x * x;
    s = x * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 2.75573192e-6f, -1.98412698e-4f), 8.33333333e-3f),
                          -1.66666667e-1f), 1.0f);
    c = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, -2.75573192e-7f, 2.48015873e-5f), -1.38888889e-3f),
                               4.16666667e-2f), -0.5f), 1.0f);
}

// sin and cos of a phase of u >= 0 quarter turns, reduced in FP64.
__device__ __forceinline__ void sincos_quarter(double u, float& sn, float& cs) {
    const double q = rint(u);
    const float x = (float)((u - q) * (0.5 * kPi));
    const int iq = (int)q & 3;
    float s, c;
    sincos_poly(x, s, c);
    const bool swap = iq & 1;
    const float sv = swap ? c : s, cv = swap ? s : c;
    sn = (iq & 2) ? -sv : sv;
    cs = ((iq + 1) & 2) ? -cv : cv;
}

// e^y for y <= 0 from an FP64 argument: y = n ln2 + r, |r| <= ln2/2, e^r by its
// Taylor series to r^8 (truncation < 6e-9), scaled by 2^n.
__device__ __forceinline__ float exp_neg(double y) {
    const double n = rint(y * 1.4426950408889634);
    const float r = (float)(y - n * 0.6931471805599453);
    float p = fmaf(r, 2.48015873e-5f, 1.98412698e-4f);
    p = fmaf(r, p, 1.38888889e-3f);
    p = fmaf(r, p, 8.33333333e-3f);
    p = fmaf(r, p, 4.16666667e-2f);
    p = fmaf(r, p, 1.66666667e-1f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    return ldexpf(p, (int)n);
}

struct C2 {
    double r, i;
};

// Moments of one triangle pair for one wavenumber ([..][0] real, [..][1] imag).
struct Mom {
    double m0[2], mr[3][2], ms[3][2], mrs[2], nx[3][2], nd[3][2];
};

__device__ __forceinline__ void mom_zero(Mom& m) {
    m.m0[0] = m.m0[1] = m.mrs[0] = m.mrs[1] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++)
        m.mr[d][0] = m.mr[d][1] = m.ms[d][0] = m.ms[d][1] = m.nx[d][0] = m.nx[d][1] = m.nd[d][0] = m.nd[d][1] = 0.0;
}

// Inner sums over source points j for one test point i:
// a = sum w_j G, b = sum w_j G r_j, d = sum w_j g, c = sum w_j g r_j.
struct Inner {
    double a[2], b[3][2], d[2], c[3][2];
};

__device__ __forceinline__ void inner_zero(Inner& s) {
    s.a[0] = s.a[1] = s.d[0] = s.d[1] = 0.0;
#pragma unroll
    for (int e = 0; e < 3; e++) s.b[e][0] = s.b[e][1] = s.c[e][0] = s.c[e][1] = 0.0;
}

__device__ __forceinline__ void inner_add(Inner& s, double w, double Gr, double Gi, double gr, double gi,
                                          const double4& q) {
    const double wGr = w * Gr, wGi = w * Gi, wgr = w * gr, wgi = w * gi;
    s.a[0] += wGr;
    s.a[1] += wGi;
    s.d[0] += wgr;
    s.d[1] += wgi;
    s.b[0][0] = fma(wGr, q.x, s.b[0][0]); s.b[0][1] = fma(wGi, q.x, s.b[0][1]);
    s.b[1][0] = fma(wGr, q.y, s.b[1][0]); s.b[1][1] = fma(wGi, q.y, s.b[1][1]);
    s.b[2][0] = fma(wGr, q.z, s.b[2][0]); s.b[2][1] = fma(wGi, q.z, s.b[2][1]);
    s.c[0][0] = fma(wgr, q.x, s.c[0][0]); s.c[0][1] = fma(wgi, q.x, s.c[0][1]);
    s.c[1][0] = fma(wgr, q.y, s.c[1][0]); s.c[1][1] = fma(wgi, q.y, s.c[1][1]);
    s.c[2][0] = fma(wgr, q.z, s.c[2][0]); s.c[2][1] = fma(wgi, q.z, s.c[2][1]);
}

// Folds the inner sums of test point p (weight p.w) into the moments.
__device__ __forceinline__ void mom_fold(Mom& m, const Inner& s, const double4& p) {
    const double w = p.w;
    const double r[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int c = 0; c < 2; c++) {
        m.m0[c] = fma(w, s.a[c], m.m0[c]);
        double dotb = 0.0;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            m.mr[e][c] = fma(w * r[e], s.a[c], m.mr[e][c]);
            m.ms[e][c] = fma(w, s.b[e][c], m.ms[e][c]);
            dotb = fma(r[e], s.b[e][c], dotb);
            m.nd[e][c] = fma(w, r[e] * s.d[c] - s.c[e][c], m.nd[e][c]);
        }
        m.mrs[c] = fma(w, dotb, m.mrs[c]);
        m.nx[0][c] = fma(w, r[1] * s.c[2][c] - r[2] * s.c[1][c], m.nx[0][c]);
        m.nx[1][c] = fma(w, r[2] * s.c[0][c] - r[0] * s.c[2][c], m.nx[1][c]);
        m.nx[2][c] = fma(w, r[0] * s.c[1][c] - r[1] * s.c[0][c], m.nx[2][c]);
    }
}

// lm = Mrs - p_a.Ms - p_b.Mr + (p_a.p_b) M0 and km = (p_b - p_a).Nx +
// (p_b x p_a).Nd of one half pair (formulas in the file header).
__device__ __forceinline__ void raw_parts(const Mom& m, const double* pa, const double* pb, double lm[2],
                                          double km[2]) {
    const double papb = pa[0] * pb[0] + pa[1] * pb[1] + pa[2] * pb[2];
    const double cr[3] = {pb[1] * pa[2] - pb[2] * pa[1], pb[2] * pa[0] - pb[0] * pa[2],
                          pb[0] * pa[1] - pb[1] * pa[0]};  // p_b x p_a
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double v = m.mrs[c] + papb * m.m0[c];
        double k = 0.0;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            v -= pa[e] * m.ms[e][c] + pb[e] * m.mr[e][c];
            k += (pb[e] - pa[e]) * m.nx[e][c] + cr[e] * m.nd[e][c];
        }
        lm[c] = v;
        km[c] = k;
    }
}

// L_ab and K_ab from the moments (formulas in the file header).
__device__ __forceinline__ void half_pair(const Mom& m, double kr, double ki, double ca, const double* pa, double cb,
                                          const double* pb, C2& L, C2& K) {
    double lm[2], km[2];
    raw_parts(m, pa, pb, lm, km);
    // ik * lm - (4i/k) m0, with i/k = (ki + i kr)/|k|^2
    const double k2 = kr * kr + ki * ki;
    const double fr = 4.0 * ki / k2, fi = 4.0 * kr / k2;
    const double Lr = -ki * lm[0] - kr * lm[1] - (fr * m.m0[0] - fi * m.m0[1]);
    const double Li = -ki * lm[1] + kr * lm[0] - (fr * m.m0[1] + fi * m.m0[0]);
    const double cc = ca * cb;
    L.r = cc * Lr;
    L.i = cc * Li;
    K.r = cc * km[0];
    K.i = cc * km[1];
}

// Writes the 9 half-pair values of pair (t, s) to their slots; `corr` (self
// pairs only) holds the symmetrised static part per wavenumber and half pair.
__device__ __forceinline__ void write_record(double* __restrict__ H, int ntile, const int* __restrict__ hrwg,
                                             const Mom* mom, const AsmParams& p, const double* __restrict__ hc,
                                             const double* __restrict__ hp, int t, int s, const C2* corr) {
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
        const double ca = hc[t * 3 + a];
        const double pa[3] = {hp[(t * 3 + a) * 3], hp[(t * 3 + a) * 3 + 1], hp[(t * 3 + a) * 3 + 2]};
#pragma unroll 1
        for (int b = 0; b < 3; b++) {
            long long dup;
            const long long slot = half_pair_slot(hrwg, hc, t, a, s, b, ntile, &dup);
            if (slot < 0) continue;
            const double cb = hc[s * 3 + b];
            const double pb[3] = {hp[(s * 3 + b) * 3], hp[(s * 3 + b) * 3 + 1], hp[(s * 3 + b) * 3 + 2]};
            C2 L[2], K[2];
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                half_pair(mom[kk], p.kr[kk], p.ki[kk], ca, pa, cb, pb, L[kk], K[kk]);
                if (corr) {
                    L[kk].r += corr[kk * 9 + a * 3 + b].r;
                    L[kk].i += corr[kk * 9 + a * 3 + b].i;
                }
            }
            store_slot(H, slot, dup,
                       make_double2(p.eta[0] * L[0].r + p.eta[1] * L[1].r, p.eta[0] * L[0].i + p.eta[1] * L[1].i),
                       make_double2(L[0].r / p.eta[0] + L[1].r / p.eta[1], L[0].i / p.eta[0] + L[1].i / p.eta[1]),
                       make_double2(K[0].r + K[1].r, K[0].i + K[1].i));
        }
    }
}

// acc += (a[0] + i a[1]) (x + i y)
__device__ __forceinline__ void cmac(double acc[2], const double a[2], double x, double y) {
    acc[0] = fma(a[0], x, fma(-a[1], y, acc[0]));
    acc[1] = fma(a[0], y, fma(a[1], x, acc[1]));
}

// Writes the 9 half-pair values of pair (t, s) combined with q (RecCoef);
// `sv` (self pairs, moments of G only) holds the static parts vec_ab (9
// values) and S_P (sv[9]), which enter U and V with c_a c_b / (4 pi).
__device__ __forceinline__ void write_record_coef(double* __restrict__ H, int ntile, const int* __restrict__ hrwg,
                                                  const Mom* mom, const RecCoef& q, const double* __restrict__ hc,
                                                  const double* __restrict__ hp, int t, int s, const double* sv) {
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
        const double ca = hc[t * 3 + a];
        const double pa[3] = {hp[(t * 3 + a) * 3], hp[(t * 3 + a) * 3 + 1], hp[(t * 3 + a) * 3 + 2]};
#pragma unroll 1
        for (int b = 0; b < 3; b++) {
            long long dup;
            const long long slot = half_pair_slot(hrwg, hc, t, a, s, b, ntile, &dup);
            if (slot < 0) continue;
            const double cb = hc[s * 3 + b];
            const double pb[3] = {hp[(s * 3 + b) * 3], hp[(s * 3 + b) * 3 + 1], hp[(s * 3 + b) * 3 + 2]};
            const double cc = ca * cb;
            double P[2] = {0.0, 0.0}, Q[2] = {0.0, 0.0}, K[2] = {0.0, 0.0};
#pragma unroll
            for (int w = 0; w < 2; w++) {
                double lm[2], km[2];
                raw_parts(mom[w], pa, pb, lm, km);
                double U[2] = {cc * lm[0], cc * lm[1]}, V[2] = {cc * mom[w].m0[0], cc * mom[w].m0[1]};
                if (sv) {
                    U[0] += cc * kInv4Pi * sv[a * 3 + b];
                    V[0] += cc * kInv4Pi * sv[9];
                }
                cmac(P, q.aP[w], U[0], U[1]);
                cmac(P, q.bP[w], V[0], V[1]);
                cmac(Q, q.aQ[w], U[0], U[1]);
                cmac(Q, q.bQ[w], V[0], V[1]);
                cmac(K, q.cK[w], cc * km[0], cc * km[1]);
            }
            store_slot(H, slot, dup, make_double2(P[0], P[1]), make_double2(Q[0], Q[1]), make_double2(K[0], K[1]));
        }
    }
}

__device__ __forceinline__ bool touching(const int4& a, const int4& b) {
    return a.x == b.x || a.x == b.y || a.x == b.z || a.y == b.x || a.y == b.y || a.y == b.z || a.z == b.x ||
           a.z == b.y || a.z == b.z;
}

// ---------------------------------------------------------------------------
// Non-touching triangle pairs.  Block = 16 x 16 triangles (t in the row tile,
// s in the column tile), only tiles with row <= column; thread (ti, tj).
// ---------------------------------------------------------------------------
// MODE 0: forward records; 1: RecCoef records of the moments of G, g;
// 2: RecCoef records of the moments of dG/dk = iR G and dg/dk = -k G.
template <int NQ, bool LOSSY, int MODE>
__global__ __launch_bounds__(256) void k_tripair_far(int T, const double4* __restrict__ qp,
                                                     const int4* __restrict__ tv, const double* __restrict__ hc,
                                                     const double* __restrict__ hp, const int* __restrict__ hrwg,
                                                     int ntile, AsmParams p, RecCoefSet q,
                                                     double* __restrict__ H) {
    const int bt = blockIdx.y, bs = blockIdx.x;
    if (bt > bs) return;
    __shared__ double4 sT[16 * NQ], sS[16 * NQ];
    __shared__ int4 vT[16], vS[16];
    for (int i = threadIdx.x; i < 16 * NQ; i += 256) {
        const int t = bt * 16 + i / NQ, s = bs * 16 + i / NQ;
        sT[i] = t < T ? qp[t * NQ + i % NQ] : make_double4(0, 0, 0, 0);
        sS[i] = s < T ? qp[s * NQ + i % NQ] : make_double4(0, 0, 0, 0);
    }
    if (threadIdx.x < 16) {
        const int t = bt * 16 + threadIdx.x, s = bs * 16 + threadIdx.x;
        vT[threadIdx.x] = t < T ? tv[t] : make_int4(-1, -1, -1, 0);
        vS[threadIdx.x] = s < T ? tv[s] : make_int4(-2, -2, -2, 0);
    }
    __syncthreads();
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
    const int t = bt * 16 + ti, s = bs * 16 + tj;
    if (t >= T || s >= T || t > s) return;
    if (touching(vT[ti], vS[tj])) return;  // k_tripair_near owns this record

    const double kq0 = p.kr[0] * (2.0 / kPi), kq1 = p.kr[1] * (2.0 / kPi);  // quarter turns per unit length
    const float kif0 = (float)p.ki[0], kif1 = (float)p.ki[1];
    const float krf0 = (float)p.kr[0], krf1 = (float)p.kr[1];
    const bool lossy0 = p.ki[0] != 0.0;
    Mom mom[2];
    mom_zero(mom[0]);
    mom_zero(mom[1]);
#pragma unroll 1
    for (int i = 0; i < NQ; i++) {
        const double4 P = sT[ti * NQ + i];
        Inner in0, in1;
        inner_zero(in0);
        inner_zero(in1);
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const double4 Q = sS[tj * NQ + j];
            const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            double y = (double)__builtin_amdgcn_rsqf((float)r2);
            y = y * fma(-0.5 * r2, y * y, 1.5);  // Newton step in FP64
            const double R = r2 * y;
            const float invR = (float)y;
            const float amp = invR * (float)kInv4Pi;
            {  // exterior wavenumber
                float sn, cs;
                sincos_quarter(R * kq0, sn, cs);
                const float a0 = lossy0 ? amp * exp_neg(-p.ki[0] * R) : amp;
                const float Gr = a0 * cs, Gi = a0 * sn;
                if (MODE == 2) {
                    const float Rf = (float)R;
                    inner_add(in0, Q.w, -Rf * Gi, Rf * Gr, -(krf0 * Gr - kif0 * Gi), -(krf0 * Gi + kif0 * Gr), Q);
                } else {
                    const float ar = (-kif0 - invR) * invR, ai = krf0 * invR;
                    inner_add(in0, Q.w, Gr, Gi, Gr * ar - Gi * ai, Gr * ai + Gi * ar, Q);
                }
            }
            {  // interior wavenumber
                float sn, cs;
                sincos_quarter(R * kq1, sn, cs);
                const float a1 = LOSSY ? amp * exp_neg(-p.ki[1] * R) : amp;
                const float Gr = a1 * cs, Gi = a1 * sn;
                if (MODE == 2) {
                    const float Rf = (float)R;
                    inner_add(in1, Q.w, -Rf * Gi, Rf * Gr, -(krf1 * Gr - kif1 * Gi), -(krf1 * Gi + kif1 * Gr), Q);
                } else {
                    const float ar = (-kif1 - invR) * invR, ai = krf1 * invR;
                    inner_add(in1, Q.w, Gr, Gi, Gr * ar - Gi * ai, Gr * ai + Gi * ar, Q);
                }
            }
        }
        mom_fold(mom[0], in0, P);
        mom_fold(mom[1], in1, P);
    }
    if (MODE == 0) {
        write_record(H, ntile, hrwg, mom, p, hc, hp, t, s, nullptr);
    } else {
#pragma unroll 1
        for (int j = 0; j < q.n; j++) write_record_coef(q.H[j], ntile, hrwg, mom, q.q[j], hc, hp, t, s, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Graglia integrals at every quadrature point of its own triangle (FP64):
// P(r) = int_T 1/|r - r'| (graglia.h:11-66) and V(r) = r P - grad W with
// W(r) = A sum_q w_q |r_q - r| differentiated by central differences of step
// 1e-6 (graglia.h:69-112).  One thread per (triangle, point).
// ---------------------------------------------------------------------------
__device__ double pot_integral(const double r[3], const double v[3][3]) {
    const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn < 1e-15) return 0.0;
    n[0] *= 1.0 / nn;
    n[1] *= 1.0 / nn;
    n[2] *= 1.0 / nn;
    double res = 0.0;
    for (int i = 0; i < 3; i++) {
        const double* a = v[i];
        const double* b = v[(i + 1) % 3];
        double t[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double le = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (le < 1e-15) continue;
        t[0] *= 1.0 / le;
        t[1] *= 1.0 / le;
        t[2] *= 1.0 / le;
        const double m[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
        const double ra[3] = {r[0] - a[0], r[1] - a[1], r[2] - a[2]};
        const double d = ra[0] * m[0] + ra[1] * m[1] + ra[2] * m[2];
        const double h = ra[0] * n[0] + ra[1] * n[1] + ra[2] * n[2];
        const double bj[3] = {b[0] - r[0], b[1] - r[1], b[2] - r[2]};
        const double ai[3] = {a[0] - r[0], a[1] - r[1], a[2] - r[2]};
        const double sp = bj[0] * t[0] + bj[1] * t[1] + bj[2] * t[2];
        const double sm = ai[0] * t[0] + ai[1] * t[1] + ai[2] * t[2];
        const double Rp = sqrt(bj[0] * bj[0] + bj[1] * bj[1] + bj[2] * bj[2]);
        const double Rm = sqrt(ai[0] * ai[0] + ai[1] * ai[1] + ai[2] * ai[2]);
        if (Rp + sp > 1e-15 && Rm + sm > 1e-15) {
            const double la = (Rp + sp) / (Rm + sm);
            if (la > 0) res += d * log(la);
        }
        if (fabs(h) > 1e-15) {
            const double R0 = d * d + h * h;
            res -= fabs(h) * (atan2(d * sp, R0 + fabs(h) * Rp) - atan2(d * sm, R0 + fabs(h) * Rm));
        }
    }
    return res;
}

template <int NQ>
__global__ __launch_bounds__(256) void k_graglia(int T, const double4* __restrict__ qp,
                                                 const int4* __restrict__ tv, const double* __restrict__ verts,
                                                 const double* __restrict__ qw, double4* __restrict__ pv) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T * NQ) return;
    const int t = idx / NQ;
    const int4 iv = tv[t];
    const int ids[3] = {iv.x, iv.y, iv.z};
    double v[3][3];
    for (int k = 0; k < 3; k++)
        for (int e = 0; e < 3; e++) v[k][e] = verts[ids[k] * 3 + e];
    const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                 cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    const double4 q = qp[idx];
    const double r[3] = {q.x, q.y, q.z};
    const double P = pot_integral(r, v);
    const double eps = 1e-6;
    double g[3];
    for (int d = 0; d < 3; d++) {
        double rp[3] = {r[0], r[1], r[2]}, rm[3] = {r[0], r[1], r[2]};
        rp[d] += eps;
        rm[d] -= eps;
        double wp = 0.0, wm = 0.0;
        for (int j = 0; j < NQ; j++) {
            const double4 s = qp[t * NQ + j];
            const double ap[3] = {s.x - rp[0], s.y - rp[1], s.z - rp[2]};
            const double am[3] = {s.x - rm[0], s.y - rm[1], s.z - rm[2]};
            wp += qw[j] * sqrt(ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2]);
            wm += qw[j] * sqrt(am[0] * am[0] + am[1] * am[1] + am[2] * am[2]);
        }
        g[d] = (area * wp - area * wm) / (2.0 * eps);
    }
    pv[idx] = make_double4(P, r[0] * P - g[0], r[1] * P - g[1], r[2] * P - g[2]);
}

// ---------------------------------------------------------------------------
// Touching triangle pairs (shared vertex, shared edge, or the same triangle),
// all in FP64.  Self pairs use G - 1/(4 pi R) with the R = 0 value
// i Re(k)/(4 pi) of the reference (assembly.cu:63-77), skip K (assembly.cu:94),
// and add the symmetrised Graglia static part (assembly.cu:290-317):
//   C_ab = c_a c_b / (4 pi) [ -(4i/k) S_P + ik (S_rV - (p_a+p_b).(S_Pr+S_V)/2
//                                                 + (p_a.p_b) S_P) ]
// with S_P = sum jw P, S_Pr = sum jw P r, S_V = sum jw V, S_rV = sum jw r.V.
// ---------------------------------------------------------------------------
template <int NQ, int MODE>
__global__ __launch_bounds__(128) void k_tripair_near(int T, int n_near, const int2* __restrict__ pairs,
                                                      const double4* __restrict__ qp,
                                                      const double* __restrict__ hc,
                                                      const double* __restrict__ hp,
                                                      const double4* __restrict__ pv,
                                                      const int* __restrict__ hrwg, int ntile, AsmParams p,
                                                      RecCoefSet q, double* __restrict__ H) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_near) return;
    const int t = pairs[idx].x, s = pairs[idx].y;
    const bool self = t == s;
    // In MODE 2 the R = 0 value of a self pair depends on the coefficient
    // set, so such a pair forms its moments once per set.
    const int nmom = (MODE == 2 && self) ? q.n : 1;
#pragma unroll 1
    for (int js = 0; js < nmom; js++) {
        Mom mom[2];
        mom_zero(mom[0]);
        mom_zero(mom[1]);
#pragma unroll 1
        for (int i = 0; i < NQ; i++) {
            const double4 P = qp[t * NQ + i];
            Inner in[2];
            inner_zero(in[0]);
            inner_zero(in[1]);
#pragma unroll 1
            for (int j = 0; j < NQ; j++) {
                const double4 Q = qp[s * NQ + j];
                const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double R = sqrt(r2);
#pragma unroll
                for (int kk = 0; kk < 2; kk++) {
                    const double kr = p.kr[kk], ki = p.ki[kk];
                    double Gr, Gi, gr = 0.0, gi = 0.0;
                    if (self && R < 1e-12) {
                        if (MODE == 2) {
                            Gr = q.q[js].r0[kk][0];
                            Gi = q.q[js].r0[kk][1];
                        } else {
                            Gr = 0.0;
                            Gi = kr * kInv4Pi;
                        }
                    } else {
                        double sn, cs;
                        sincos(kr * R, &sn, &cs);
                        const double amp = exp(-ki * R) * kInv4Pi / R;
                        Gr = amp * cs;
                        Gi = amp * sn;
                        const bool with_g = !self && R > 1e-6;
                        if (MODE == 2) {  // dG/dk = iR G (the static 1/(4 pi R) has none), dg/dk = -k G
                            if (with_g) {
                                gr = -(kr * Gr - ki * Gi);
                                gi = -(kr * Gi + ki * Gr);
                            }
                            const double t0 = -R * Gi;
                            Gi = R * Gr;
                            Gr = t0;
                        } else {
                            if (with_g) {
                                const double ar = (-ki - 1.0 / R) / R, ai = kr / R;
                                gr = Gr * ar - Gi * ai;
                                gi = Gr * ai + Gi * ar;
                            }
                            if (self) Gr -= kInv4Pi / R;
                        }
                    }
                    inner_add(in[kk], Q.w, Gr, Gi, gr, gi, Q);
                }
            }
            mom_fold(mom[0], in[0], P);
            mom_fold(mom[1], in[1], P);
        }
        if (!self || MODE == 2) {
            if (MODE == 0) {
                write_record(H, ntile, hrwg, mom, p, hc, hp, t, s, nullptr);
            } else {
                const int j0 = nmom > 1 ? js : 0, j1 = nmom > 1 ? js + 1 : q.n;
#pragma unroll 1
                for (int j = j0; j < j1; j++)
                    write_record_coef(q.H[j], ntile, hrwg, mom, q.q[j], hc, hp, t, s, nullptr);
            }
            continue;
        }
        double SP = 0.0, SPr[3] = {0, 0, 0}, SV[3] = {0, 0, 0}, SrV = 0.0;
        for (int iq = 0; iq < NQ; iq++) {
            const double4 x = qp[t * NQ + iq];
            const double4 w = pv[t * NQ + iq];  // (P, Vx, Vy, Vz)
            SP += x.w * w.x;
            SPr[0] += x.w * w.x * x.x;
            SPr[1] += x.w * w.x * x.y;
            SPr[2] += x.w * w.x * x.z;
            SV[0] += x.w * w.y;
            SV[1] += x.w * w.z;
            SV[2] += x.w * w.w;
            SrV += x.w * (x.x * w.y + x.y * w.z + x.z * w.w);
        }
        C2 corr[18];
        double sv[10];
        sv[9] = SP;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const double ca = hc[t * 3 + a], cb = hc[t * 3 + b];
                const double* pa = hp + (t * 3 + a) * 3;
                const double* pb = hp + (t * 3 + b) * 3;
                double sym = 0.0, papb = 0.0;
                for (int e = 0; e < 3; e++) {
                    sym += (pa[e] + pb[e]) * (SPr[e] + SV[e]);
                    papb += pa[e] * pb[e];
                }
                const double vec = SrV - 0.5 * sym + papb * SP;  // real
                sv[a * 3 + b] = vec;
                for (int kk = 0; kk < 2; kk++) {
                    const double kr = p.kr[kk], ki = p.ki[kk], k2 = kr * kr + ki * ki;
                    // -(4i/k) SP + ik vec, i/k = (ki + i kr)/|k|^2, ik = -ki + i kr
                    const double re = -4.0 * ki / k2 * SP - ki * vec;
                    const double im = -4.0 * kr / k2 * SP + kr * vec;
                    const double f = ca * cb * kInv4Pi;
                    corr[kk * 9 + a * 3 + b].r = f * re;
                    corr[kk * 9 + a * 3 + b].i = f * im;
                }
            }
        if (MODE == 0) {
            write_record(H, ntile, hrwg, mom, p, hc, hp, t, t, corr);
        } else {
#pragma unroll 1
            for (int j = 0; j < q.n; j++) write_record_coef(q.H[j], ntile, hrwg, mom, q.q[j], hc, hp, t, t, sv);
        }
    }
}

// ---------------------------------------------------------------------------
// Z[m][n] from the four half-pair slots of RWGs m and n, PMCHWT blocks:
// Z[m][n] = P, Z[m][N+n] = -Ks, Z[N+m][n] = Ks, Z[N+m][N+n] = Q.
// The slots are shared by (m, n) and (n, m) (P, Q and Ks are symmetric in
// the RWG pair), so one read of them gives both: block = 16 x 16 tile of
// (m, n) with the m tile at or before the n tile, thread (i, j) sums the pair
// (tm0 + i, tn0 + j) when m <= n (its 4 slots: 192 contiguous bytes, the
// block's 48 KB contiguous) and writes its row segment; the (n, m) entries go
// out as row segments of the mirrored tile through LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(int N, int ld, const double* __restrict__ H,
                                                double* __restrict__ Zr, double* __restrict__ Zi,
                                                float* __restrict__ Wr, float* __restrict__ Wi) {
    if (blockIdx.y > blockIdx.x) return;
    __shared__ double sv[6][kGt][kGt + 1];  // (P, Q, Ks) x (re, im) of pair (i, j)
    const int tm0 = blockIdx.y * kGt, tn0 = blockIdx.x * kGt;
    const int i = threadIdx.x / kGt, j = threadIdx.x % kGt;
    const int m = tm0 + i, n = tn0 + j;
    const bool own = m < N && n < N && m <= n;
    double P0 = 0, P1 = 0, Q0 = 0, Q1 = 0, K0 = 0, K1 = 0;
    if (own) {
        // slots in the order (hm, hn) = (+, +), (+, -), (-, +), (-, -)
        const double2* h = reinterpret_cast<const double2*>(H + slot_index(m, n, 0, 0, gridDim.x) * kSlot);
        double2 v[12];
#pragma unroll
        for (int k = 0; k < 12; k++) v[k] = h[k];
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
            P0 += v[3 * sl].x; P1 += v[3 * sl].y;
            Q0 += v[3 * sl + 1].x; Q1 += v[3 * sl + 1].y;
            K0 += v[3 * sl + 2].x; K1 += v[3 * sl + 2].y;
        }
        const size_t r0 = (size_t)m * ld, r1 = (size_t)(N + m) * ld;
        if (Zr) {
            Zr[r0 + n] = P0;      Zi[r0 + n] = P1;
            Zr[r0 + N + n] = -K0; Zi[r0 + N + n] = -K1;
            Zr[r1 + n] = K0;      Zi[r1 + n] = K1;
            Zr[r1 + N + n] = Q0;  Zi[r1 + N + n] = Q1;
        }
        if (Wr) {
            Wr[r0 + n] = (float)P0;      Wi[r0 + n] = (float)P1;
            Wr[r0 + N + n] = (float)-K0; Wi[r0 + N + n] = (float)-K1;
            Wr[r1 + n] = (float)K0;      Wi[r1 + n] = (float)K1;
            Wr[r1 + N + n] = (float)Q0;  Wi[r1 + N + n] = (float)Q1;
        }
    }
    sv[0][i][j] = P0; sv[1][i][j] = P1;
    sv[2][i][j] = Q0; sv[3][i][j] = Q1;
    sv[4][i][j] = K0; sv[5][i][j] = K1;
    __syncthreads();
    // mirrored entry: row n = tn0 + i, column m = tm0 + j, from pair (j, i)
    const int mm = tm0 + j, nn = tn0 + i;
    if (mm < N && nn < N && mm < nn) {
        const double p0 = sv[0][j][i], p1 = sv[1][j][i], q0 = sv[2][j][i], q1 = sv[3][j][i];
        const double k0 = sv[4][j][i], k1 = sv[5][j][i];
        const size_t r0 = (size_t)nn * ld, r1 = (size_t)(N + nn) * ld;
        if (Zr) {
            Zr[r0 + mm] = p0;      Zi[r0 + mm] = p1;
            Zr[r0 + N + mm] = -k0; Zi[r0 + N + mm] = -k1;
            Zr[r1 + mm] = k0;      Zi[r1 + mm] = k1;
            Zr[r1 + N + mm] = q0;  Zi[r1 + N + mm] = q1;
        }
        if (Wr) {
            Wr[r0 + mm] = (float)p0;      Wi[r0 + mm] = (float)p1;
            Wr[r0 + N + mm] = (float)-k0; Wi[r0 + N + mm] = (float)-k1;
            Wr[r1 + mm] = (float)k0;      Wi[r1 + mm] = (float)k1;
            Wr[r1 + N + mm] = (float)q0;  Wi[r1 + N + mm] = (float)q1;
        }
    }
}

// Identity on rows/columns [n, ld): keeps the padded system block-diagonal.
template <class F>
__global__ void k_pad(int n, int ld, F* __restrict__ Ar, F* __restrict__ Ai) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= ld) return;
    if (i < n && j < n) return;
    Ar[(size_t)i * ld + j] = (i == j) ? F(1) : F(0);
    Ai[(size_t)i * ld + j] = F(0);
}

// <Z_H, W> over the 2N x 2N matrix whose entries k_gather would form from the
// records, for the NS record arrays H.p[j] at once.  The records give Z the
// symmetry Z[n][m] = Z[m][n] within the P and Q blocks and Ks[n][m] =
// Ks[m][n], so one read of the four half-pair records of (m, n), m < n,
// serves both (m, n) and (n, m):
//   <Z, W> = sum_{m <= n} P Ws00 - Ks Ws01 + Ks Ws10 + Q Ws11,
//   Ws_b = W_b[m][n] + W_b[n][m] (m < n), W_b[m][m] (m = n),
// with W_b the four N x N blocks of W.  Block = 16 x 16 tile of (m, n) with
// the m tile at or before the n tile; W_b[n][m] comes through LDS so both W
// reads are row segments.  Each block sums in a fixed tree order into
// part[(tile * NS + j) * 2 + {0, 1}], tile = blockIdx.y * gridDim.x +
// blockIdx.x (tiles below the diagonal write zeros).
struct RecArrays {
    const double* p[kMaxRecSets];
};

constexpr int kCt = kGt;  // contraction tile edge (RWGs): the slot tiles

template <int NS>
__global__ __launch_bounds__(256) void k_contract(int N, RecArrays H, const double* __restrict__ Wr,
                                                  const double* __restrict__ Wi, long ldw,
                                                  double* __restrict__ part) {
    __shared__ double wt[8][kCt][kCt + 1];  // W_b[n][m], b = 0..3, (re, im)
    __shared__ double red[2 * NS][256];
    const int tm0 = blockIdx.y * kCt, tn0 = blockIdx.x * kCt;
    const long tile = (long)blockIdx.y * gridDim.x + blockIdx.x;
    if (blockIdx.y > blockIdx.x) {
        if (threadIdx.x < 2 * NS) part[tile * 2 * NS + threadIdx.x] = 0.0;
        return;
    }
    const int i = threadIdx.x / kCt, j = threadIdx.x % kCt;
    {  // W_b[n][m] for n = tn0 + i, m = tm0 + j, row segments
        const int n = tn0 + i, m = tm0 + j;
        const bool ok = n < N && m < N;
        const size_t rn = (size_t)n * ldw, rN = (size_t)(N + n) * ldw;
        wt[0][i][j] = ok ? Wr[rn + m] : 0.0;
        wt[1][i][j] = ok ? Wi[rn + m] : 0.0;
        wt[2][i][j] = ok ? Wr[rn + N + m] : 0.0;
        wt[3][i][j] = ok ? Wi[rn + N + m] : 0.0;
        wt[4][i][j] = ok ? Wr[rN + m] : 0.0;
        wt[5][i][j] = ok ? Wi[rN + m] : 0.0;
        wt[6][i][j] = ok ? Wr[rN + N + m] : 0.0;
        wt[7][i][j] = ok ? Wi[rN + N + m] : 0.0;
    }
    __syncthreads();
    const int m = tm0 + i, n = tn0 + j;
    double acc[NS][2];
#pragma unroll
    for (int q = 0; q < NS; q++) acc[q][0] = acc[q][1] = 0.0;
    if (m < N && n < N && m <= n) {
        double P[NS][2], Q[NS][2], K[NS][2];
#pragma unroll
        for (int q = 0; q < NS; q++) P[q][0] = P[q][1] = Q[q][0] = Q[q][1] = K[q][0] = K[q][1] = 0.0;
        const long long s0 = slot_index(m, n, 0, 0, gridDim.x) * kSlot;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            const double2* h = reinterpret_cast<const double2*>(H.p[q] + s0);
#pragma unroll
            for (int sl = 0; sl < 4; sl++) {  // (hm, hn) = (+, +), (+, -), (-, +), (-, -)
                const double2 a = h[3 * sl], b = h[3 * sl + 1], c = h[3 * sl + 2];
                P[q][0] += a.x; P[q][1] += a.y;
                Q[q][0] += b.x; Q[q][1] += b.y;
                K[q][0] += c.x; K[q][1] += c.y;
            }
        }
        const size_t r0 = (size_t)m * ldw, r1 = (size_t)(N + m) * ldw;
        const double f = m < n ? 1.0 : 0.0;  // the diagonal counts once
        const double w00r = Wr[r0 + n] + f * wt[0][j][i], w00i = Wi[r0 + n] + f * wt[1][j][i];
        const double w01r = Wr[r0 + N + n] + f * wt[2][j][i], w01i = Wi[r0 + N + n] + f * wt[3][j][i];
        const double w10r = Wr[r1 + n] + f * wt[4][j][i], w10i = Wi[r1 + n] + f * wt[5][j][i];
        const double w11r = Wr[r1 + N + n] + f * wt[6][j][i], w11i = Wi[r1 + N + n] + f * wt[7][j][i];
#pragma unroll
        for (int q = 0; q < NS; q++) {
            const double mK[2] = {-K[q][0], -K[q][1]};
            cmac(acc[q], P[q], w00r, w00i);
            cmac(acc[q], mK, w01r, w01i);
            cmac(acc[q], K[q], w10r, w10i);
            cmac(acc[q], Q[q], w11r, w11i);
        }
    }
#pragma unroll
    for (int q = 0; q < NS; q++) {
        red[2 * q][threadIdx.x] = acc[q][0];
        red[2 * q + 1][threadIdx.x] = acc[q][1];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 2 * NS; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2 * NS) part[tile * 2 * NS + threadIdx.x] = red[threadIdx.x][0];
}

// out[c] = sum over i < np of part[i * ncomp + c], block c, fixed order.
__global__ __launch_bounds__(1024) void k_sum_parts(long np, int ncomp, const double* __restrict__ part,
                                                    double* __restrict__ out) {
    __shared__ double red[1024];
    const int c = blockIdx.x;
    double a = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024) a += part[i * ncomp + c];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// Graglia data (modes 0 and 1) and the records of all triangle pairs into c->H.
template <int NQ, int MODE>
void launch_records(bem_ctx* c, const AsmParams& p, const RecCoefSet& q, bool lossy, const double* dverts,
                    const double* dqw) {
    const DevMesh& d = c->dm;
    const int T = d.T;
    if (MODE != 2)
        hipLaunchKernelGGL(k_graglia<NQ>, dim3((T * NQ + 255) / 256), dim3(256), 0, c->stream, T, d.qp, d.tv,
                           dverts, dqw, d.pv);
    const int nb = (T + 15) / 16;
    const int ntile = (d.N + kGt - 1) / kGt;
    const double npair = (double)T * (T + 1) / 2;
    int kt = kt_mark(c, KS_TRIPAIR_FAR);
    if (lossy)
        hipLaunchKernelGGL((k_tripair_far<NQ, true, MODE>), dim3(nb, nb), dim3(256), 0, c->stream, T, d.qp, d.tv,
                           d.hc, d.hp, d.hrwg, ntile, p, q, c->H);
    else
        hipLaunchKernelGGL((k_tripair_far<NQ, false, MODE>), dim3(nb, nb), dim3(256), 0, c->stream, T, d.qp, d.tv,
                           d.hc, d.hp, d.hrwg, ntile, p, q, c->H);
    kt_close(c, kt, npair * NQ * NQ, npair * kRec * 8.0);  // unit: triangle point pairs
    kt = kt_mark(c, KS_TRIPAIR_NEAR);
    if (d.n_near > 0)
        hipLaunchKernelGGL((k_tripair_near<NQ, MODE>), dim3((d.n_near + 127) / 128), dim3(128), 0, c->stream, T,
                           d.n_near, d.near, d.qp, d.hc, d.hp, d.pv, d.hrwg, ntile, p, q, c->H);
    kt_close(c, kt, (double)d.n_near * NQ * NQ, (double)d.n_near * kRec * 8.0);
}

template <int NQ>
int launch_nq(bem_ctx* c, const AsmParams& p, bool lossy, double* Zr, double* Zi, float* Wr, float* Wi,
              const double* dverts, const double* dqw) {
    const DevMesh& d = c->dm;
    launch_records<NQ, 0>(c, p, RecCoefSet{}, lossy, dverts, dqw);
    const int N = d.N;
    const int kt = kt_mark(c, KS_GATHER);
    const int gt = (N + kGt - 1) / kGt;
    hipLaunchKernelGGL(k_gather, dim3(gt, gt), dim3(256), 0, c->stream, N, c->npad, c->H, Zr, Zi, Wr, Wi);
    // the slots of each unordered RWG pair read once, every Z entry written once
    kt_close(c, kt, 0.0, (double)N * N * (2.0 * 48.0 + (Zr ? 64.0 : 0.0) + (Wr ? 32.0 : 0.0)));
    const int n = 2 * N;
    if (c->npad > n) {
        if (Zr)
            hipLaunchKernelGGL(k_pad<double>, dim3((c->npad + 255) / 256, c->npad), dim3(256), 0, c->stream, n,
                               c->npad, Zr, Zi);
        if (Wr)
            hipLaunchKernelGGL(k_pad<float>, dim3((c->npad + 255) / 256, c->npad), dim3(256), 0, c->stream, n,
                               c->npad, Wr, Wi);
    }
    return 0;
}

}  // namespace

// Device copies of the vertex coordinates and quadrature weights used by
// k_graglia (context scratch).
static int asm_prepare_graglia(bem_ctx* c, const double** dverts, const double** dqw) {
    const int nv = c->mesh.nv();
    const size_t sbytes = (size_t)(nv * 3 + 32) * sizeof(double);
    void* s = nullptr;
    if (int rc = ctx_scratch(c, sbytes, &s)) return rc;
    std::vector<double> host(nv * 3 + 32, 0.0);
    for (int i = 0; i < nv; i++) {
        host[3 * i] = c->mesh.verts[i].x;
        host[3 * i + 1] = c->mesh.verts[i].y;
        host[3 * i + 2] = c->mesh.verts[i].z;
    }
    for (int q = 0; q < c->quad.npts; q++) host[nv * 3 + q] = c->quad.wts[q];
    BEM_HIP(c, hipMemcpy(s, host.data(), sbytes, hipMemcpyHostToDevice));
    *dverts = static_cast<const double*>(s);
    *dqw = *dverts + nv * 3;
    return 0;
}

// Bytes of one slot array: 4 slots per RWG pair of every tile with m tile <=
// n tile (slot_index).
static size_t slot_array_bytes(int N) {
    const size_t nt = (size_t)((N + kGt - 1) / kGt);
    return nt * (nt + 1) / 2 * kGt * kGt * 4 * kSlot * sizeof(double);
}

// Slot buffer and the k_graglia inputs.
static int asm_prepare(bem_ctx* c, const double** dverts, const double** dqw) {
    const size_t need = slot_array_bytes(c->dm.N);
    if (c->H_bytes < need) {
        if (c->H) (void)hipFree(c->H);
        c->H = nullptr;
        c->H_bytes = 0;
        if (hipMalloc(&c->H, need) != hipSuccess) {
            c->H = nullptr;
            return bem_fail(c, BEM_ERR_NOMEM, "assembly scratch: cannot allocate " + std::to_string(need) + " bytes");
        }
        c->H_bytes = need;
    }
    return asm_prepare_graglia(c, dverts, dqw);
}

// ---------------------------------------------------------------------------
// Near-field singular corrections of the fast (FMM / pFFT) operators
// (BemFmmOperator::precompute_corrections, bem_fmm.cu:506-686), all in FP64.
// For the halves a, b of two RWG functions on one triangle and wavenumber k:
//   C_L = ik sum_ij f_a(r_i).f_b(r_j) DG_ij jw_i jw_j
//         - (i/k) d_a d_b sum_ij DG_ij jw_i jw_j
//         + ik/(4 pi) sum_i f_a(r_i).c_b (V_i - p_b P_i) jw_i
//         - (i/k) d_a d_b / (4 pi) sum_i P_i jw_i,
//   DG = -1/(4 pi R) (R > 1e-12), ik/(4 pi) at R = 0 (the exact limit of the
//   smooth kernel, where the dense path keeps the clamped i Re(k)/(4 pi));
//   C_K = -sum_{R > 1e-12} G (ik - 1/R)/R f_a(r_i).((r_i - r_j) x f_b(r_j)) jw_i jw_j,
// which vanishes on a flat triangle up to rounding (the triple product of
// three in-plane vectors).  One thread per (triangle, half pair, wavenumber);
// its sums run in the reference's order.  out[(t * 9 + a * 3 + b) * 2 + w] =
// (C_L, minus C_K) as four doubles.
// ---------------------------------------------------------------------------
template <int NQ>
__global__ __launch_bounds__(256) void k_near_corr(int T, const double4* __restrict__ qp,
                                                   const double* __restrict__ hc, const double* __restrict__ hp,
                                                   const int* __restrict__ hrwg, const double4* __restrict__ pv,
                                                   double ker, double kei, double kir, double kii,
                                                   double* __restrict__ out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T * 18) return;
    const int t = idx / 18, r = idx % 18, pr = r / 2, w = r % 2, a = pr / 3, b = pr % 3;
    double* o = out + (long)idx * 4;
    if (hrwg[t * 3 + a] < 0 || hrwg[t * 3 + b] < 0) {
        o[0] = o[1] = o[2] = o[3] = 0.0;
        return;
    }
    const double kr = w ? kir : ker, ki = w ? kii : kei;
    const double ca = hc[t * 3 + a], cb = hc[t * 3 + b];
    const double pa[3] = {hp[(t * 3 + a) * 3], hp[(t * 3 + a) * 3 + 1], hp[(t * 3 + a) * 3 + 2]};
    const double pb[3] = {hp[(t * 3 + b) * 3], hp[(t * 3 + b) * 3 + 1], hp[(t * 3 + b) * 3 + 2]};
    const double da = 2.0 * ca, db = 2.0 * cb;
    // ik = (-ki, kr); i/k = (ki + i kr) / |k|^2
    const double ikr = -ki, iki = kr;
    const double k2 = kr * kr + ki * ki, iokr = ki / k2, ioki = kr / k2;
    double dvr = 0, dvi = 0, dsr = 0, dsi = 0, kcr = 0, kci = 0;
    for (int i = 0; i < NQ; i++) {
        const double4 qi = qp[t * NQ + i];
        const double fa[3] = {ca * (qi.x - pa[0]), ca * (qi.y - pa[1]), ca * (qi.z - pa[2])};
        for (int j = 0; j < NQ; j++) {
            const double4 qj = qp[t * NQ + j];
            const double dx = qi.x - qj.x, dy = qi.y - qj.y, dz = qi.z - qj.z;
            const double R = sqrt(dx * dx + dy * dy + dz * dz);
            const double jp = qi.w * qj.w;
            const double fb[3] = {cb * (qj.x - pb[0]), cb * (qj.y - pb[1]), cb * (qj.z - pb[2])};
            const double fd = fa[0] * fb[0] + fa[1] * fb[1] + fa[2] * fb[2];
            double gr, gi;  // DG
            if (R > 1e-12) {
                gr = -1.0 / (4.0 * kPi * R);
                gi = 0.0;
            } else {
                gr = ikr / (4.0 * kPi);
                gi = iki / (4.0 * kPi);
            }
            // dv += ik * fd * DG * jp; ds += -iok * da * db * DG * jp
            const double tr = ikr * fd, ti = iki * fd;
            dvr += (tr * gr - ti * gi) * jp;
            dvi += (tr * gi + ti * gr) * jp;
            const double sr = -iokr * da * db, si = -ioki * da * db;
            dsr += (sr * gr - si * gi) * jp;
            dsi += (sr * gi + si * gr) * jp;
            if (R >= 1e-12) {
                // G = exp(ikR) / (4 pi R), exp(ikR) = exp(-ki R) (cos kr R + i sin kr R)
                double sn, cs;
                sincos(kr * R, &sn, &cs);
                const double e = exp(-ki * R) / (4.0 * kPi * R);
                const double Gr = e * cs, Gi = e * sn;
                const double hr = (ikr - 1.0 / R) / R, hi = iki / R;  // (ik - 1/R) / R
                const double ggr = Gr * hr - Gi * hi, ggi = Gr * hi + Gi * hr;
                const double cx = dy * fb[2] - dz * fb[1], cy = dz * fb[0] - dx * fb[2], cz = dx * fb[1] - dy * fb[0];
                const double tp = (fa[0] * cx + fa[1] * cy + fa[2] * cz) * jp;
                kcr += ggr * tp;
                kci += ggi * tp;
            }
        }
    }
    double avr = 0, avi = 0, asr = 0, asi = 0;
    for (int i = 0; i < NQ; i++) {
        const double4 qi = qp[t * NQ + i];
        const double4 g = pv[t * NQ + i];  // P, V
        const double fa[3] = {ca * (qi.x - pa[0]), ca * (qi.y - pa[1]), ca * (qi.z - pa[2])};
        const double fo[3] = {cb * (g.y - pb[0] * g.x), cb * (g.z - pb[1] * g.x), cb * (g.w - pb[2] * g.x)};
        const double fdo = fa[0] * fo[0] + fa[1] * fo[1] + fa[2] * fo[2];
        avr += ikr * fdo * qi.w * kInv4Pi;
        avi += iki * fdo * qi.w * kInv4Pi;
        asr += -iokr * da * db * g.x * qi.w * kInv4Pi;
        asi += -ioki * da * db * g.x * qi.w * kInv4Pi;
    }
    o[0] = ((dvr + dsr) + avr) + asr;
    o[1] = ((dvi + dsi) + avi) + asi;
    o[2] = kcr;
    o[3] = kci;
}

// CSR values from the per-triangle contributions: entry e takes the slots
// src[2e] and src[2e + 1] (-1: none) in that (triangle) order.
__global__ __launch_bounds__(256) void k_near_gather(int nnz, const int* __restrict__ src,
                                                     const double* __restrict__ contrib, double* __restrict__ vals) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    for (int w = 0; w < 2; w++) {
        double lr = 0, li = 0, kr = 0, ki = 0;
        for (int u = 0; u < 2; u++) {
            const int sl = src[2 * e + u];
            if (sl < 0) continue;
            const double* c = contrib + ((long)sl * 2 + w) * 4;
            lr += c[0];
            li += c[1];
            kr -= c[2];
            ki -= c[3];
        }
        // vals: L_ext, K_ext, L_int, K_int, each nnz interleaved complex
        double* L = vals + (long)(2 * w) * 2 * nnz;
        double* K = vals + (long)(2 * w + 1) * 2 * nnz;
        L[2 * e] = lr;
        L[2 * e + 1] = li;
        K[2 * e] = kr;
        K[2 * e + 1] = ki;
    }
}

int asm_launch(bem_ctx* c, const AsmParams& p, bool want64, bool want32) {
    const double *dverts, *dqw;
    if (int rc = asm_prepare(c, &dverts, &dqw)) return rc;
    const bool lossy = p.ki[1] != 0.0;
    double* Zr = want64 ? c->Zr : nullptr;
    double* Zi = want64 ? c->Zi : nullptr;
    float* Wr = want32 ? c->Wr : nullptr;
    float* Wi = want32 ? c->Wi : nullptr;
    int rc;
    switch (c->dm.Nq) {
        case 4: rc = launch_nq<4>(c, p, lossy, Zr, Zi, Wr, Wi, dverts, dqw); break;
        case 13: rc = launch_nq<13>(c, p, lossy, Zr, Zi, Wr, Wi, dverts, dqw); break;
        default: rc = launch_nq<7>(c, p, lossy, Zr, Zi, Wr, Wi, dverts, dqw); break;
    }
    if (rc) return rc;
    BEM_HIP(c, hipGetLastError());
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int near_corr_launch(bem_ctx* c, double ker, double kei, double kir, double kii, int* row_ptr, int* col_idx,
                     double* const out[4]) {
    const DevMesh& d = c->dm;
    const int T = d.T, N = d.N;
    // CSR pattern (bem_fmm.cu:540-567) and, per entry, its contributing
    // (triangle, half a, half b) slots in triangle order
    std::vector<int> hr((size_t)T * 3);
    BEM_HIP(c, hipMemcpy(hr.data(), d.hrwg, hr.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<std::vector<std::pair<int, int>>> rows(N);  // (column, slot)
    for (int t = 0; t < T; t++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const int m = hr[t * 3 + a], n = hr[t * 3 + b];
                if (m >= 0 && n >= 0) rows[m].push_back({n, t * 9 + a * 3 + b});
            }
    std::vector<int> rp(N + 1, 0), ci, src;
    for (int m = 0; m < N; m++) {
        std::stable_sort(rows[m].begin(), rows[m].end(),
                         [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return x.first < y.first; });
        for (size_t i = 0; i < rows[m].size();) {
            size_t j = i;
            int s0 = -1, s1 = -1;
            while (j < rows[m].size() && rows[m][j].first == rows[m][i].first) {
                if (s0 < 0)
                    s0 = rows[m][j].second;
                else if (s1 < 0)
                    s1 = rows[m][j].second;
                else
                    return bem_fail(c, BEM_ERR_MESH, "near corrections: a basis pair shares more than two triangles");
                j++;
            }
            ci.push_back(rows[m][i].first);
            src.push_back(s0);
            src.push_back(s1);
            i = j;
        }
        rp[m + 1] = (int)ci.size();
    }
    const int nnz = rp[N];
    if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
    if (!col_idx) return nnz;
    std::copy(ci.begin(), ci.end(), col_idx);
    if (nnz == 0) return 0;
    const double *dverts, *dqw;
    if (int rc = asm_prepare_graglia(c, &dverts, &dqw)) return rc;
    // device: contributions (T * 18 * 4), sources (2 nnz ints), values (8 nnz)
    const size_t bc = (size_t)T * 18 * 4 * sizeof(double), bs = ((size_t)2 * nnz * sizeof(int) + 255) / 256 * 256;
    const size_t bv = (size_t)8 * nnz * sizeof(double);
    char* w = nullptr;
    if (hipMalloc(&w, bc + bs + bv) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "near corrections: work area");
    double* contrib = reinterpret_cast<double*>(w);
    int* dsrc = reinterpret_cast<int*>(w + bc);
    double* dvals = reinterpret_cast<double*>(w + bc + bs);
    hipError_t e = hipMemcpyAsync(dsrc, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const dim3 g1((unsigned)((T * 18 + 255) / 256)), g2((unsigned)((nnz + 255) / 256));
        switch (d.Nq) {
            case 4:
                hipLaunchKernelGGL(k_graglia<4>, dim3((T * 4 + 255) / 256), dim3(256), 0, c->stream, T, d.qp, d.tv,
                                   dverts, dqw, d.pv);
                hipLaunchKernelGGL(k_near_corr<4>, g1, dim3(256), 0, c->stream, T, d.qp, d.hc, d.hp, d.hrwg, d.pv, ker,
                                   kei, kir, kii, contrib);
                break;
            case 13:
                hipLaunchKernelGGL(k_graglia<13>, dim3((T * 13 + 255) / 256), dim3(256), 0, c->stream, T, d.qp, d.tv,
                                   dverts, dqw, d.pv);
                hipLaunchKernelGGL(k_near_corr<13>, g1, dim3(256), 0, c->stream, T, d.qp, d.hc, d.hp, d.hrwg, d.pv,
                                   ker, kei, kir, kii, contrib);
                break;
            default:
                hipLaunchKernelGGL(k_graglia<7>, dim3((T * 7 + 255) / 256), dim3(256), 0, c->stream, T, d.qp, d.tv,
                                   dverts, dqw, d.pv);
                hipLaunchKernelGGL(k_near_corr<7>, g1, dim3(256), 0, c->stream, T, d.qp, d.hc, d.hp, d.hrwg, d.pv, ker,
                                   kei, kir, kii, contrib);
                break;
        }
        hipLaunchKernelGGL(k_near_gather, g2, dim3(256), 0, c->stream, nnz, dsrc, contrib, dvals);
        e = hipGetLastError();
    }
    std::vector<double> hv((size_t)8 * nnz);
    if (e == hipSuccess) e = hipMemcpyAsync(hv.data(), dvals, bv, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(w);
    if (e != hipSuccess) return bem_fail(c, BEM_ERR_HIP, std::string("near corrections: ") + hipGetErrorString(e));
    for (int a = 0; a < 4; a++)
        if (out[a]) std::copy(hv.begin() + (size_t)a * 2 * nnz, hv.begin() + (size_t)(a + 1) * 2 * nnz, out[a]);
    return nnz;
}

namespace {
// out[2e], out[2e+1] = Z64[rows[e]][cols[e]] (planar, ld = npad)
__global__ __launch_bounds__(256) void k_z_entries(int count, const int* __restrict__ rows,
                                                   const int* __restrict__ cols, const double* __restrict__ Zr,
                                                   const double* __restrict__ Zi, long ld, double* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= count) return;
    const long o = (long)rows[e] * ld + cols[e];
    out[2 * e] = Zr[o];
    out[2 * e + 1] = Zi[o];
}
}  // namespace

int z_entries_launch(bem_ctx* c, const int* rows, const int* cols, int count, double* out) {
    hipLaunchKernelGGL(k_z_entries, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c->stream, count, rows, cols,
                       c->Zr, c->Zi, (long)c->npad, out);
    BEM_HIP(c, hipGetLastError());
    return 0;
}

size_t asm_record_bytes(const bem_ctx* c) { return slot_array_bytes(c->dm.N); }

int asm_contract_launch(bem_ctx* c, const AsmParams& p, const RecCoef* q, int nq, double* extra, const double* Wr,
                        const double* Wi, long ldw, double* out) {
    if (nq < 1 || nq > kMaxRecSets) return bem_fail(c, BEM_ERR_ARG, "contraction: 1..3 coefficient sets");
    if (nq > 1 && !extra) return bem_fail(c, BEM_ERR_ARG, "contraction: no record arrays for the extra sets");
    for (int j = 1; j < nq; j++)
        if (q[j].deriv != q[0].deriv) return bem_fail(c, BEM_ERR_ARG, "contraction: sets of one moment kind");
    const double *dverts, *dqw;
    if (int rc = asm_prepare(c, &dverts, &dqw)) return rc;
    RecCoefSet qs{};
    qs.n = nq;
    const size_t rec_doubles = asm_record_bytes(c) / sizeof(double);
    RecArrays ra{};
    for (int j = 0; j < nq; j++) {
        qs.q[j] = q[j];
        qs.H[j] = j == 0 ? c->H : extra + (size_t)(j - 1) * rec_doubles;
        ra.p[j] = qs.H[j];
    }
    const bool lossy = p.ki[1] != 0.0;
    const int Nq = c->dm.Nq;
    if (q[0].deriv) {
        if (Nq == 4) launch_records<4, 2>(c, p, qs, lossy, dverts, dqw);
        else if (Nq == 13) launch_records<13, 2>(c, p, qs, lossy, dverts, dqw);
        else launch_records<7, 2>(c, p, qs, lossy, dverts, dqw);
    } else {
        if (Nq == 4) launch_records<4, 1>(c, p, qs, lossy, dverts, dqw);
        else if (Nq == 13) launch_records<13, 1>(c, p, qs, lossy, dverts, dqw);
        else launch_records<7, 1>(c, p, qs, lossy, dverts, dqw);
    }
    const int N = c->dm.N;
    const int gx = (N + kCt - 1) / kCt;
    const long np = (long)gx * gx;
    const int ncomp = 2 * nq;
    double* part = nullptr;
    if (hipMalloc(&part, (np + 1) * ncomp * sizeof(double)) != hipSuccess)
        return bem_fail(c, BEM_ERR_NOMEM, "contraction partial sums");
    const dim3 grid(gx, gx);
    if (nq == 1)
        hipLaunchKernelGGL(k_contract<1>, grid, dim3(256), 0, c->stream, N, ra, Wr, Wi, ldw, part);
    else if (nq == 2)
        hipLaunchKernelGGL(k_contract<2>, grid, dim3(256), 0, c->stream, N, ra, Wr, Wi, ldw, part);
    else
        hipLaunchKernelGGL(k_contract<3>, grid, dim3(256), 0, c->stream, N, ra, Wr, Wi, ldw, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(ncomp), dim3(1024), 0, c->stream, np, ncomp, part, part + np * ncomp);
    const hipError_t e1 =
        hipMemcpyAsync(out, part + np * ncomp, ncomp * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(part);
    if (e1 != hipSuccess || e2 != hipSuccess)
        return bem_fail(c, BEM_ERR_HIP, std::string("contraction: ") + hipGetErrorString(e1 ? e1 : e2));
    return 0;
}
