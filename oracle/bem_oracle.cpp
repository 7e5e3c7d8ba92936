// bem_oracle.cpp -- FP64 CPU restatement of the reference's dense PMCHWT path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
// entry of __graft_entry__.py and the cpu_baseline leg of bench.py may load
// the library built from it.  No product code (the HIP library, the host C++
// library, the CLI) links or calls it.
//
// Every function restates one piece of the reference, in double precision,
// following the cited file:line of KirillSalnikov/BEM-CPP.  The reference GPU
// assembly computes in FP32 (assembly.cu:38-112); this restatement keeps its
// algorithm (same passes, same quadrature, same singular extraction, same
// symmetrization) and evaluates it in FP64, which is the semantics of the
// reference's FP64 twin python/bem_core.py:787-1025.
//
// Documented choice (R = 0 in the same-triangle smooth kernel): both reference
// implementations clamp R (assembly.cu:63 to 1e-7f in FP32, bem_core.py:892 to
// 1e-15 in FP64).  At the configured wavenumbers both evaluate
// G_smooth(0) = i*Re(k)/(4*pi) (the FP32 exp(-Im k*1e-7) and the FP64
// exp(-Im k*1e-15) both round to exactly 1).  The oracle uses that value.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;
static const double PI = 3.14159265358979323846;
static const double INV4PI = 1.0 / (4.0 * PI);  // types.h:49

struct V3 {
    double x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
};

// ---------------------------------------------------------------------------
// Quadrature: Dunavant rules, quadrature.h:14-70 (constants as written there).
// ---------------------------------------------------------------------------
struct Quad {
    int n;
    double l1[25], l2[25], w[25];
};

static Quad quad_rule(int order) {
    Quad q;
    if (order == 4) {  // quadrature.h:15-23
        q.n = 4;
        double p[4][2] = {{1.0 / 3, 1.0 / 3}, {0.6, 0.2}, {0.2, 0.6}, {0.2, 0.2}};
        double w[4] = {-27.0 / 48, 25.0 / 48, 25.0 / 48, 25.0 / 48};
        for (int i = 0; i < 4; i++) { q.l1[i] = p[i][0]; q.l2[i] = p[i][1]; q.w[i] = w[i]; }
    } else if (order == 13) {  // quadrature.h:38-65
        q.n = 13;
        double p[13][2] = {
            {1.0 / 3, 1.0 / 3},
            {0.260345966079038, 0.260345966079038}, {0.260345966079038, 0.479308067841923},
            {0.479308067841923, 0.260345966079038}, {0.065130102902216, 0.065130102902216},
            {0.065130102902216, 0.869739794195568}, {0.869739794195568, 0.065130102902216},
            {0.048690315425316, 0.312865496004875}, {0.312865496004875, 0.048690315425316},
            {0.638444188569809, 0.312865496004875}, {0.312865496004875, 0.638444188569809},
            {0.048690315425316, 0.638444188569809}, {0.638444188569809, 0.048690315425316}};
        double w[13] = {-0.149570044467670,
                        0.175615257433204, 0.175615257433204, 0.175615257433204,
                        0.053347235608839, 0.053347235608839, 0.053347235608839,
                        0.077113760890257, 0.077113760890257, 0.077113760890257,
                        0.077113760890257, 0.077113760890257, 0.077113760890257};
        for (int i = 0; i < 13; i++) { q.l1[i] = p[i][0]; q.l2[i] = p[i][1]; q.w[i] = w[i]; }
    } else {  // order 7 and the fallback for any other order, quadrature.h:24-37, 66-68
        q.n = 7;
        double a1 = 0.059715871789770, a2 = 0.470142064105115;
        double b1 = 0.797426985353087, b2 = 0.101286507323456;
        double p[7][2] = {{1.0 / 3, 1.0 / 3}, {a1, a2}, {a2, a1}, {a2, a2}, {b1, b2}, {b2, b1}, {b2, b2}};
        double w[7] = {0.225, 0.13239415278851, 0.13239415278851, 0.13239415278851,
                       0.12593918054483, 0.12593918054483, 0.12593918054483};
        for (int i = 0; i < 7; i++) { q.l1[i] = p[i][0]; q.l2[i] = p[i][1]; q.w[i] = w[i]; }
    }
    return q;
}

// ---------------------------------------------------------------------------
// Meshes: mesh.cpp:11-154.
// ---------------------------------------------------------------------------
struct MeshD {
    std::vector<V3> v;
    std::vector<int> t;
    int nt() const { return (int)t.size() / 3; }
    void tri(int i, V3& a, V3& b, V3& c) const { a = v[t[3 * i]]; b = v[t[3 * i + 1]]; c = v[t[3 * i + 2]]; }
    double area(int i) const { V3 a, b, c; tri(i, a, b, c); return 0.5 * (b - a).cross(c - a).norm(); }
};

static MeshD icosphere(double radius, int refinements) {  // mesh.cpp:11-67
    double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<V3> v = {{-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0},
                         {0, -1, phi}, {0, 1, phi}, {0, -1, -phi}, {0, 1, -phi},
                         {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
    double n0 = v[0].norm();
    for (auto& p : v) p = p * (1.0 / n0);
    std::vector<int> t = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
                          1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                          3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
                          4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int r = 0; r < refinements; r++) {
        std::map<std::pair<int, int>, int> mid;
        std::vector<int> nt;
        auto get = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V3 m = (v[a] + v[b]) * 0.5;
            double nn = m.norm();
            m = V3(m.x / nn, m.y / nn, m.z / nn);
            int idx = (int)v.size();
            v.push_back(m);
            mid[key] = idx;
            return idx;
        };
        int ntri = (int)t.size() / 3;
        for (int i = 0; i < ntri; i++) {
            int a = t[3 * i], b = t[3 * i + 1], c = t[3 * i + 2];
            int ab = get(a, b), bc = get(b, c), ca = get(c, a);
            int s[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nt.insert(nt.end(), s, s + 12);
        }
        t = nt;
    }
    for (auto& p : v) p = p * radius;
    MeshD m;
    m.v = v;
    m.t = t;
    return m;
}

static MeshD subdivide_flat(const MeshD& in) {  // mesh.cpp:70-100
    std::map<std::pair<int, int>, int> mid;
    std::vector<V3> v = in.v;
    std::vector<int> nt;
    auto get = [&](int a, int b) {
        auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        V3 m = (v[a] + v[b]) * 0.5;
        int idx = (int)v.size();
        v.push_back(m);
        mid[key] = idx;
        return idx;
    };
    int ntri = in.nt();
    for (int i = 0; i < ntri; i++) {
        int a = in.t[3 * i], b = in.t[3 * i + 1], c = in.t[3 * i + 2];
        int ab = get(a, b), bc = get(b, c), ca = get(c, a);
        int s[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
        nt.insert(nt.end(), s, s + 12);
    }
    MeshD m;
    m.v = v;
    m.t = nt;
    return m;
}

static MeshD hex_prism(double ar, int refinements) {  // mesh.cpp:102-154
    double a = std::pow(8.0 * PI / (27.0 * ar), 1.0 / 3.0);
    double H = ar * a * std::sqrt(3.0);
    std::vector<V3> v(14);
    v[0] = V3(0, 0, H / 2);
    v[7] = V3(0, 0, -H / 2);
    for (int i = 0; i < 6; i++) {
        double ang = 2.0 * PI * i / 6.0;
        double x = a * std::cos(ang), y = a * std::sin(ang);
        v[1 + i] = V3(x, y, H / 2);
        v[8 + i] = V3(x, y, -H / 2);
    }
    std::vector<int> t;
    for (int i = 0; i < 6; i++) {
        int i1 = 1 + i, i2 = 1 + (i + 1) % 6, j1 = 8 + i, j2 = 8 + (i + 1) % 6;
        int s[12] = {0, i1, i2, 7, j2, j1, i1, j1, j2, i1, j2, i2};
        t.insert(t.end(), s, s + 12);
    }
    MeshD m;
    m.v = v;
    m.t = t;
    for (int r = 0; r < refinements; r++) m = subdivide_flat(m);
    return m;
}

// ---------------------------------------------------------------------------
// RWG basis: rwg.cpp:5-43 (std::map edge order; plus = first triangle seen).
// ---------------------------------------------------------------------------
struct Rwg {
    int N = 0;
    std::vector<int> tp, tm;
    std::vector<V3> fp, fm;
    std::vector<double> len, ap, am;
};

static Rwg build_rwg(const MeshD& m) {
    std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> et;
    int ntri = m.nt();
    for (int ti = 0; ti < ntri; ti++)
        for (int j = 0; j < 3; j++) {
            int a = m.t[3 * ti + j], b = m.t[3 * ti + (j + 1) % 3], o = m.t[3 * ti + (j + 2) % 3];
            et[std::make_pair(std::min(a, b), std::max(a, b))].push_back(std::make_pair(ti, o));
        }
    Rwg r;
    for (auto& kv : et) {
        if (kv.second.size() != 2) continue;
        int tp = kv.second[0].first, vp = kv.second[0].second;
        int tm = kv.second[1].first, vm = kv.second[1].second;
        r.tp.push_back(tp);
        r.tm.push_back(tm);
        r.fp.push_back(m.v[vp]);
        r.fm.push_back(m.v[vm]);
        r.len.push_back((m.v[kv.first.second] - m.v[kv.first.first]).norm());
        r.ap.push_back(m.area(tp));
        r.am.push_back(m.area(tm));
    }
    r.N = (int)r.tp.size();
    return r;
}

// ---------------------------------------------------------------------------
// Graglia integrals: graglia.h:11-112.
// ---------------------------------------------------------------------------
static double pot_integral(const V3& r, const V3& v0, const V3& v1, const V3& v2) {  // graglia.h:11-66
    V3 vs[3] = {v0, v1, v2};
    V3 n = (v1 - v0).cross(v2 - v0);
    double nn = n.norm();
    if (nn < 1e-15) return 0.0;
    V3 nh = n * (1.0 / nn);
    double res = 0.0;
    for (int i = 0; i < 3; i++) {
        V3 vi = vs[i], vj = vs[(i + 1) % 3];
        V3 e = vj - vi;
        double le = e.norm();
        if (le < 1e-15) continue;
        V3 th = e * (1.0 / le);
        V3 mh = nh.cross(th);
        V3 di = r - vi;
        double d = di.dot(mh), h = di.dot(nh);
        V3 dj = vj - r, dii = vi - r;
        double sp = dj.dot(th), sm = dii.dot(th);
        double Rp = dj.norm(), Rm = dii.norm();
        if (Rp + sp > 1e-15 && Rm + sm > 1e-15) {
            double la = (Rp + sp) / (Rm + sm);
            if (la > 0) res += d * std::log(la);
        }
        if (std::fabs(h) > 1e-15) {
            double R0 = d * d + h * h;
            double t1 = std::atan2(d * sp, R0 + std::fabs(h) * Rp);
            double t2 = std::atan2(d * sm, R0 + std::fabs(h) * Rm);
            res -= std::fabs(h) * (t1 - t2);
        }
    }
    return res;
}

static double integral_R(const V3& r, const V3& v0, const V3& v1, const V3& v2, const Quad& q) {  // graglia.h:69-85
    double area = 0.5 * (v1 - v0).cross(v2 - v0).norm();
    double s = 0.0;
    for (int i = 0; i < q.n; i++) {
        double l0 = 1.0 - q.l1[i] - q.l2[i];
        V3 rr = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
        s += q.w[i] * (rr - r).norm();
    }
    return area * s;
}

static V3 vec_pot_integral(const V3& r, const V3& v0, const V3& v1, const V3& v2, const Quad& q) {  // graglia.h:90-112
    double P = pot_integral(r, v0, v1, v2);
    const double eps = 1e-6;
    double g[3];
    for (int d = 0; d < 3; d++) {
        V3 rp = r, rm = r;
        if (d == 0) { rp.x += eps; rm.x -= eps; }
        else if (d == 1) { rp.y += eps; rm.y -= eps; }
        else { rp.z += eps; rm.z -= eps; }
        g[d] = (integral_R(rp, v0, v1, v2, q) - integral_R(rm, v0, v1, v2, q)) / (2.0 * eps);
    }
    return V3(r.x * P - g[0], r.y * P - g[1], r.z * P - g[2]);
}

// ---------------------------------------------------------------------------
// Per-half quadrature data: assembly.cu:150-201 (kept in FP64 here).
// ---------------------------------------------------------------------------
struct Half {
    std::vector<V3> q, f;       // N*Nq points and basis values
    std::vector<double> jw;     // N*Nq
    std::vector<double> div;    // N
    std::vector<int> tri;       // N
};

static Half make_half(const Rwg& r, const MeshD& m, const Quad& q, int sign) {
    Half h;
    int N = r.N, Nq = q.n;
    h.q.resize(N * Nq);
    h.f.resize(N * Nq);
    h.jw.resize(N * Nq);
    h.div.resize(N);
    h.tri.resize(N);
    for (int n = 0; n < N; n++) {
        int ti = sign > 0 ? r.tp[n] : r.tm[n];
        V3 fv = sign > 0 ? r.fp[n] : r.fm[n];
        double A = sign > 0 ? r.ap[n] : r.am[n];
        double l = r.len[n];
        h.tri[n] = ti;
        h.div[n] = sign * l / A;
        V3 v0, v1, v2;
        m.tri(ti, v0, v1, v2);
        double c = sign * l / (2.0 * A);
        for (int i = 0; i < Nq; i++) {
            double l0 = 1.0 - q.l1[i] - q.l2[i];
            V3 rr = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
            h.q[n * Nq + i] = rr;
            h.f[n * Nq + i] = (rr - fv) * c;
            h.jw[n * Nq + i] = A * q.w[i];
        }
    }
    return h;
}

// R = 0 limit of the smooth kernel: 0 = i*Re(k)/(4 pi) as the reference evaluates
// it (default), 1 = the exact limit i*k/(4 pi) (used by tests to bound the
// Python reference's own clamp ambiguity).
static int g_r0_exact = 0;

// One (test half b, source half n) entry of one pass: assembly.cu:31-134 in FP64.
static void half_pair(const Half& T, const Half& S, int b, int n, int Nq, cplx k, cplx& Lout, cplx& Kout) {
    const cplx I(0.0, 1.0);
    const cplx ik = I * k;
    bool sing = T.tri[b] == S.tri[n];
    cplx Lvec = 0.0, Lscl = 0.0, Kacc = 0.0;
    for (int i = 0; i < Nq; i++) {
        const V3& tp = T.q[b * Nq + i];
        const V3& tf = T.f[b * Nq + i];
        double jwt = T.jw[b * Nq + i];
        for (int j = 0; j < Nq; j++) {
            const V3& sq = S.q[n * Nq + j];
            const V3& sf = S.f[n * Nq + j];
            V3 d = tp - sq;
            double R = d.norm();
            double jw = jwt * S.jw[n * Nq + j];
            double fdot = tf.dot(sf);
            cplx Gu, G;
            if (sing && R < 1e-12) {
                // Limit of (exp(ikR)-1)/(4 pi R) as evaluated by the reference (see header).
                Gu = g_r0_exact ? cplx(-k.imag() * INV4PI, k.real() * INV4PI) : cplx(0.0, k.real() * INV4PI);
                G = 0.0;
            } else {
                G = std::exp(ik * R) * (INV4PI / R);
                Gu = sing ? G - INV4PI / R : G;
            }
            Lvec += fdot * Gu * jw;
            Lscl += Gu * jw;
            if (!sing && R > 1e-6) {  // assembly.cu:94
                cplx gc = G * (ik - 1.0 / R) / R;
                double triple = tf.dot(d.cross(sf));  // = cross(tf,tp).sf - tf.(sq x sf), assembly.cu:102-109
                Kacc += gc * jw * triple;
            }
        }
    }
    double dp = T.div[b] * S.div[n];
    Lout = ik * Lvec - (I / k) * dp * Lscl;  // assembly.cu:117-132
    Kout = Kacc;
}

// Singular corrections: assembly.cu:208-321 in FP64.
static void singular_corrections(const Rwg& r, const MeshD& m, const Quad& q, cplx k, cplx* L) {
    const cplx I(0.0, 1.0);
    const cplx ik = I * k, ik_inv = I / k;
    int N = r.N, Nq = q.n;
    struct Ent { int idx; double div, coeff; V3 fv; int sign; };
    std::map<int, std::vector<Ent>> t2r;
    for (int n = 0; n < N; n++) {
        t2r[r.tp[n]].push_back({n, r.len[n] / r.ap[n], r.len[n] / (2.0 * r.ap[n]), r.fp[n], +1});
        t2r[r.tm[n]].push_back({n, -r.len[n] / r.am[n], r.len[n] / (2.0 * r.am[n]), r.fm[n], -1});
    }
    std::vector<std::pair<int, const std::vector<Ent>*>> tris;
    for (auto& kv : t2r) tris.push_back(std::make_pair(kv.first, &kv.second));
    // Each (m, n) pair gets contributions from the triangles both halves share; a
    // triangle-parallel loop would race on L[m*N+m], so accumulate per triangle and
    // apply serially in the reference's triangle order.
    std::vector<std::vector<std::pair<long, cplx>>> upd(tris.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (long it = 0; it < (long)tris.size(); it++) {
        int ti = tris[it].first;
        const std::vector<Ent>& es = *tris[it].second;
        V3 v0, v1, v2;
        m.tri(ti, v0, v1, v2);
        std::vector<V3> qp(Nq), V(Nq);
        std::vector<double> P(Nq);
        for (int i = 0; i < Nq; i++) {
            double l0 = 1.0 - q.l1[i] - q.l2[i];
            qp[i] = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
            P[i] = pot_integral(qp[i], v0, v1, v2);
            V[i] = vec_pot_integral(qp[i], v0, v1, v2, q);
        }
        for (const Ent& te : es) {
            double tA = te.sign > 0 ? r.ap[te.idx] : r.am[te.idx];
            std::vector<V3> tf(Nq);
            std::vector<double> tjw(Nq);
            double sb = 0.0;
            for (int i = 0; i < Nq; i++) {
                tf[i] = (qp[i] - te.fv) * (te.sign * te.coeff);
                tjw[i] = tA * q.w[i];
                sb += P[i] * tjw[i];
            }
            sb *= INV4PI;
            for (const Ent& se : es) {
                cplx Ls = -ik_inv * te.div * se.div * sb;
                double vi = 0.0;
                for (int i = 0; i < Nq; i++) {
                    V3 fo((V[i].x - se.fv.x * P[i]), (V[i].y - se.fv.y * P[i]), (V[i].z - se.fv.z * P[i]));
                    fo = fo * (se.sign * se.coeff);
                    vi += tf[i].dot(fo) * tjw[i];
                }
                vi *= INV4PI;
                upd[it].push_back(std::make_pair((long)te.idx * N + se.idx, Ls + ik * vi));
            }
        }
    }
    for (auto& u : upd)
        for (auto& p : u) L[p.first] += p.second;
}

// L, K for one wavenumber: assembly.cu:328-505 (three passes, pm added to both
// (m,n) and (n,m), corrections, then (X + X^T)/2).
static void assemble_LK(const Rwg& r, const MeshD& m, const Quad& q, cplx k, cplx* L, cplx* K) {
    int N = r.N, Nq = q.n;
    Half hp = make_half(r, m, q, +1), hm = make_half(r, m, q, -1);
    std::fill(L, L + (size_t)N * N, cplx(0.0));
    std::fill(K, K + (size_t)N * N, cplx(0.0));
    std::vector<cplx> Lc((size_t)N * N), Kc((size_t)N * N);
    const Half* tests[3] = {&hp, &hp, &hm};
    const Half* srcs[3] = {&hp, &hm, &hm};
    for (int pass = 0; pass < 3; pass++) {
        cplx* Lt = pass == 1 ? Lc.data() : L;
        cplx* Kt = pass == 1 ? Kc.data() : K;
#pragma omp parallel for schedule(dynamic, 4)
        for (int b = 0; b < N; b++)
            for (int n = 0; n < N; n++) {
                cplx lv, kv;
                half_pair(*tests[pass], *srcs[pass], b, n, Nq, k, lv, kv);
                Lt[(size_t)b * N + n] += lv;
                Kt[(size_t)b * N + n] += kv;
            }
    }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            L[(size_t)i * N + j] += Lc[(size_t)i * N + j] + Lc[(size_t)j * N + i];
            K[(size_t)i * N + j] += Kc[(size_t)i * N + j] + Kc[(size_t)j * N + i];
        }
    singular_corrections(r, m, q, k, L);
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) {
            cplx ls = (L[(size_t)i * N + j] + L[(size_t)j * N + i]) * 0.5;
            L[(size_t)i * N + j] = ls;
            L[(size_t)j * N + i] = ls;
            cplx ks = (K[(size_t)i * N + j] + K[(size_t)j * N + i]) * 0.5;
            K[(size_t)i * N + j] = ks;
            K[(size_t)j * N + i] = ks;
        }
}

// Z = [[ee L_e + ei L_i, -(K_e+K_i)], [K_e+K_i, L_e/ee + L_i/ei]]: pmchwt.cu:36-51.
static void assemble_pmchwt(const Rwg& r, const MeshD& m, const Quad& q, cplx ke, cplx ki, double ee, double ei,
                            cplx* Z) {
    int N = r.N, N2 = 2 * N;
    std::vector<cplx> Le((size_t)N * N), Ke((size_t)N * N), Li((size_t)N * N), Ki((size_t)N * N);
    assemble_LK(r, m, q, ke, Le.data(), Ke.data());
    assemble_LK(r, m, q, ki, Li.data(), Ki.data());
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            size_t a = (size_t)i * N + j;
            cplx Ks = Ke[a] + Ki[a];
            Z[(size_t)i * N2 + j] = ee * Le[a] + ei * Li[a];
            Z[(size_t)i * N2 + N + j] = -Ks;
            Z[(size_t)(N + i) * N2 + j] = Ks;
            Z[(size_t)(N + i) * N2 + N + j] = Le[a] / ee + Li[a] / ei;
        }
}

// Plane-wave RHS: rhs.cpp:6-64.
static void rhs_planewave(const Rwg& r, const MeshD& m, const Quad& q, cplx ke, double ee, const V3& E0,
                          const V3& kh, cplx* b) {
    int N = r.N;
    V3 H0 = kh.cross(E0) * (1.0 / ee);
    for (int i = 0; i < 2 * N; i++) b[i] = 0.0;
    for (int half = 0; half < 2; half++) {
        int sign = half == 0 ? 1 : -1;
        for (int n = 0; n < N; n++) {
            int ti = sign > 0 ? r.tp[n] : r.tm[n];
            V3 fv = sign > 0 ? r.fp[n] : r.fm[n];
            double A = sign > 0 ? r.ap[n] : r.am[n];
            double c = sign * r.len[n] / (2.0 * A);
            V3 v0, v1, v2;
            m.tri(ti, v0, v1, v2);
            cplx bE = 0.0, bH = 0.0;
            for (int i = 0; i < q.n; i++) {
                double l0 = 1.0 - q.l1[i] - q.l2[i];
                V3 rr = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
                V3 f = (rr - fv) * c;
                double pa = ke.real() * kh.dot(rr), pai = ke.imag() * kh.dot(rr);
                cplx ph = std::exp(cplx(-pai, pa));
                double jw = A * q.w[i];
                bE += f.dot(E0) * ph * jw;
                bH += f.dot(H0) * ph * jw;
            }
            b[n] += bE;
            b[N + n] += bH;
        }
    }
}

// Far field of one solution at ndir directions: farfield.cu:9-52 (FFCache) and
// farfield.cu:351-420 (compute_far_field_vec_batch_cpu).
static void farfield_vec(const Rwg& r, const MeshD& m, const Quad& q, const cplx* cJ, const cplx* cM, cplx ke,
                         double ee, const V3* rh, int ndir, cplx* Fv) {
    int N = r.N, Nq = q.n;
    std::vector<cplx> Jt(ndir * 3, 0.0), Mt(ndir * 3, 0.0);
    double kr = ke.real(), ki = ke.imag();
    for (int half = 0; half < 2; half++) {
        int sign = half == 0 ? 1 : -1;
        for (int n = 0; n < N; n++) {
            int ti = sign > 0 ? r.tp[n] : r.tm[n];
            V3 fv = sign > 0 ? r.fp[n] : r.fm[n];
            double A = sign > 0 ? r.ap[n] : r.am[n];
            double c = sign * r.len[n] / (2.0 * A);
            V3 v0, v1, v2;
            m.tri(ti, v0, v1, v2);
            std::vector<cplx> integ(ndir * 3, 0.0);
            for (int i = 0; i < Nq; i++) {
                double l0 = 1.0 - q.l1[i] - q.l2[i];
                V3 rr = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
                V3 f = (rr - fv) * c;
                double w = A * q.w[i];
                for (int d = 0; d < ndir; d++) {
                    double rd = rh[d].dot(rr);
                    double ea = std::exp(ki * rd);
                    double cr = std::cos(kr * rd) * ea, si = -std::sin(kr * rd) * ea;
                    cplx pw(cr * w, si * w);
                    integ[d * 3] += f.x * pw;
                    integ[d * 3 + 1] += f.y * pw;
                    integ[d * 3 + 2] += f.z * pw;
                }
            }
            for (int d = 0; d < ndir; d++)
                for (int c3 = 0; c3 < 3; c3++) {
                    Jt[d * 3 + c3] += integ[d * 3 + c3] * cJ[n];
                    Mt[d * 3 + c3] += integ[d * 3 + c3] * cM[n];
                }
        }
    }
    cplx pf = cplx(0.0, -1.0) * ke / (4.0 * PI);
    for (int d = 0; d < ndir; d++) {
        V3 h = rh[d];
        cplx jx = Jt[d * 3], jy = Jt[d * 3 + 1], jz = Jt[d * 3 + 2];
        cplx mx = Mt[d * 3], my = Mt[d * 3 + 1], mz = Mt[d * 3 + 2];
        cplx rj = h.x * jx + h.y * jy + h.z * jz;
        cplx px = jx - h.x * rj, py = jy - h.y * rj, pz = jz - h.z * rj;
        cplx cx = h.y * mz - h.z * my, cy = h.z * mx - h.x * mz, cz = h.x * my - h.y * mx;
        Fv[d * 3] = pf * (ee * px - cx);
        Fv[d * 3 + 1] = pf * (ee * py - cy);
        Fv[d * 3 + 2] = pf * (ee * pz - cz);
    }
}

// Bohren-Huffman 3.16 as written at farfield.cu:464-508.
static void mueller(const cplx* S1, const cplx* S2, const cplx* S3, const cplx* S4, int nt, double* M) {
    for (int t = 0; t < nt; t++) {
        double a1 = std::norm(S1[t]), a2 = std::norm(S2[t]), a3 = std::norm(S3[t]), a4 = std::norm(S4[t]);
        cplx s23 = S2[t] * std::conj(S3[t]), s14 = S1[t] * std::conj(S4[t]);
        cplx s24 = S2[t] * std::conj(S4[t]), s13 = S1[t] * std::conj(S3[t]);
        cplx s12 = S1[t] * std::conj(S2[t]), s34 = S3[t] * std::conj(S4[t]);
        cplx s21 = S2[t] * std::conj(S1[t]), s42 = S4[t] * std::conj(S2[t]), s43 = S4[t] * std::conj(S3[t]);
#define MM(i, j) M[((i)*4 + (j)) * nt + t]
        MM(0, 0) = 0.5 * (a1 + a2 + a3 + a4);
        MM(0, 1) = 0.5 * (a2 - a1 + a4 - a3);
        MM(0, 2) = s23.real() + s14.real();
        MM(0, 3) = s23.imag() - s14.imag();
        MM(1, 0) = 0.5 * (a2 - a1 - a4 + a3);
        MM(1, 1) = 0.5 * (a2 + a1 - a4 - a3);
        MM(1, 2) = s23.real() - s14.real();
        MM(1, 3) = s23.imag() + s14.imag();
        MM(2, 0) = s24.real() + s13.real();
        MM(2, 1) = s24.real() - s13.real();
        MM(2, 2) = s12.real() + s34.real();
        MM(2, 3) = s21.imag() + s43.imag();
        MM(3, 0) = s42.imag() + s13.imag();
        MM(3, 1) = s42.imag() - s13.imag();
        MM(3, 2) = s12.imag() - s34.imag();
        MM(3, 3) = s12.real() - s34.real();
#undef MM
    }
}

// ---------------------------------------------------------------------------
// Orientation quadrature: orient.cpp:4-163.
// ---------------------------------------------------------------------------
struct M3 { double m[3][3]; };

static M3 euler(double al, double be, double ga) {  // orient.cpp:4-36
    double ca = std::cos(al), sa = std::sin(al), cb = std::cos(be), sb = std::sin(be);
    double cg = std::cos(ga), sg = std::sin(ga);
    double A[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    double B[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    double C[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
    M3 R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            R.m[i][j] = 0;
            for (int k = 0; k < 3; k++) {
                double t = 0;
                for (int l = 0; l < 3; l++) t += B[k][l] * C[l][j];
                R.m[i][j] += A[i][k] * t;
            }
        }
    return R;
}

static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {  // orient.cpp:39-70
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; i++) {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int j = 2; j <= n; j++) {
                double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = -p1 / dp;
            z += dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        x[i] = z;
        double p0 = 1.0, p1 = z;
        for (int j = 2; j <= n; j++) {
            double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (z * p1 - p0) / (z * z - 1.0);
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

struct Orient { M3 RT; double w; };

static std::vector<Orient> gen_orient(int na, int nb, int ng, int bsym, int gsym, bool mirror) {  // orient.cpp:73-124
    std::vector<double> mu, mw;
    if (bsym == 2) {
        std::vector<double> fx, fw;
        gauss_legendre(2 * nb, fx, fw);
        mu.resize(nb);
        mw.resize(nb);
        int j = 0;
        for (int i = 0; i < 2 * nb; i++)
            if (fx[i] > 0) { mu[j] = fx[i]; mw[j] = fw[i] * 2.0; j++; }
    } else {
        gauss_legendre(nb, mu, mw);
    }
    double da = 2.0 * PI / na, dg = 2.0 * PI / ng;
    int eg = mirror ? gsym * 2 : gsym;
    double gstep = 2.0 * PI / (eg * ng);
    std::vector<Orient> out;
    for (int ia = 0; ia < na; ia++)
        for (int ib = 0; ib < nb; ib++)
            for (int ig = 0; ig < ng; ig++) {
                double al = ia * da, be = std::acos(mu[ib]);
                double ga = mirror ? (ig + 0.5) * gstep : ig * gstep;
                M3 R = euler(al, be, ga);
                Orient o;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) o.RT.m[i][j] = R.m[j][i];
                o.w = da * mw[ib] * dg / (8.0 * PI * PI);
                out.push_back(o);
            }
    return out;
}

static void sort_orient(std::vector<Orient>& o) {  // orient.cpp:127-163
    int n = (int)o.size();
    if (n <= 2) return;
    std::vector<bool> used(n, false);
    std::vector<Orient> s;
    s.push_back(o[0]);
    used[0] = true;
    for (int step = 1; step < n; step++) {
        const M3& P = s.back().RT;
        int best = -1;
        double bd = 1e30;
        for (int j = 0; j < n; j++) {
            if (used[j]) continue;
            double d = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double df = P.m[a][b] - o[j].RT.m[a][b];
                    d += df * df;
                }
            if (d < bd) { bd = d; best = j; }
        }
        s.push_back(o[best]);
        used[best] = true;
    }
    o = s;
}

static V3 mul(const M3& A, const V3& v) {
    return V3(A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z, A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
              A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z);
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting, replacing cuSOLVER Zgetrf/Zgetrs
// (solver.cu:7-142).  Z is row-major n x n (overwritten by P Z = L U); B is
// column-major n x nrhs and is overwritten by X with Z X = B.  Returns the
// LAPACK-style info (0, or i+1 for the first exactly singular pivot).
// ---------------------------------------------------------------------------
static int lu_factor_rowmajor(cplx* A, int n, std::vector<int>& piv) {
    piv.resize(n);
    int info = 0;
    const int nb = 64;
    for (int kb = 0; kb < n; kb += nb) {
        int ke = std::min(n, kb + nb);
        for (int k = kb; k < ke; k++) {  // unblocked panel, updates limited to panel columns
            int p = k;
            double best = std::abs(A[(size_t)k * n + k]);
            for (int i = k + 1; i < n; i++) {
                double a = std::abs(A[(size_t)i * n + k]);
                if (a > best) { best = a; p = i; }
            }
            piv[k] = p;
            if (p != k)
                for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            cplx d = A[(size_t)k * n + k];
            if (d == cplx(0.0)) {
                if (!info) info = k + 1;
                continue;
            }
            cplx inv = 1.0 / d;
#pragma omp parallel for schedule(static) if (n - k > 512)
            for (int i = k + 1; i < n; i++) {
                cplx l = A[(size_t)i * n + k] * inv;
                A[(size_t)i * n + k] = l;
                for (int j = k + 1; j < ke; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            }
        }
        if (ke >= n) break;
        // U12 = L11^{-1} A12 (rows kb..ke, columns ke..n)
        for (int i = kb + 1; i < ke; i++) {
            cplx* ri = A + (size_t)i * n;
            for (int l = kb; l < i; l++) {
                cplx f = ri[l];
                const cplx* rl = A + (size_t)l * n;
                for (int j = ke; j < n; j++) ri[j] -= f * rl[j];
            }
        }
        // A22 -= L21 U12, parallel over rows, real arithmetic for vectorisation
#pragma omp parallel for schedule(static)
        for (int i = ke; i < n; i++) {
            double* ri = reinterpret_cast<double*>(A + (size_t)i * n);
            for (int l = kb; l < ke; l++) {
                double fr = ri[2 * l], fi = ri[2 * l + 1];
                const double* rl = reinterpret_cast<const double*>(A + (size_t)l * n);
                for (int j = ke; j < n; j++) {
                    double ur = rl[2 * j], ui = rl[2 * j + 1];
                    ri[2 * j] -= fr * ur - fi * ui;
                    ri[2 * j + 1] -= fr * ui + fi * ur;
                }
            }
        }
    }
    return info;
}

static void lu_solve_rowmajor(const cplx* A, int n, const std::vector<int>& piv, cplx* B, int nrhs) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nrhs; c++) {
        cplx* x = B + (size_t)c * n;
        for (int k = 0; k < n; k++)
            if (piv[k] != k) std::swap(x[k], x[piv[k]]);
        for (int i = 0; i < n; i++) {
            cplx s = x[i];
            const cplx* ri = A + (size_t)i * n;
            for (int l = 0; l < i; l++) s -= ri[l] * x[l];
            x[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            cplx s = x[i];
            const cplx* ri = A + (size_t)i * n;
            for (int l = i + 1; l < n; l++) s -= ri[l] * x[l];
            x[i] = s / ri[i];
        }
    }
}

// ---------------------------------------------------------------------------
// JSON writer: output.cpp:5-64.
// ---------------------------------------------------------------------------
static int write_json(const char* fn, const double* M, const double* th, int nt, double ka, double nre, double nim,
                      int ref, int na, int nb, int ng, double ta, double ts, double tf, double tt) {
    FILE* f = std::fopen(fn, "w");
    if (!f) return -1;
    std::fprintf(f, "{\n");
    std::fprintf(f, "  \"ka\": %.6f,\n", ka);
    std::fprintf(f, "  \"ri\": [%.6f, %.6f],\n", nre, nim);
    std::fprintf(f, "  \"refinements\": %d,\n", ref);
    std::fprintf(f, "  \"orient\": [%d, %d, %d],\n", na, nb, ng);
    std::fprintf(f, "  \"ntheta\": %d,\n", nt);
    std::fprintf(f, "  \"timing\": {\n");
    std::fprintf(f, "    \"assembly_s\": %.2f,\n", ta);
    std::fprintf(f, "    \"solve_s\": %.2f,\n", ts);
    std::fprintf(f, "    \"farfield_s\": %.2f,\n", tf);
    std::fprintf(f, "    \"total_s\": %.2f\n", tt);
    std::fprintf(f, "  },\n");
    std::fprintf(f, "  \"theta\": [");
    for (int t = 0; t < nt; t++) {
        std::fprintf(f, "%.4f", th[t] * 180.0 / PI);
        if (t < nt - 1) std::fprintf(f, ", ");
    }
    std::fprintf(f, "],\n");
    std::fprintf(f, "  \"mueller\": [\n");
    for (int i = 0; i < 4; i++) {
        std::fprintf(f, "    [\n");
        for (int j = 0; j < 4; j++) {
            std::fprintf(f, "      [");
            for (int t = 0; t < nt; t++) {
                std::fprintf(f, "%.8e", M[(i * 4 + j) * nt + t]);
                if (t < nt - 1) std::fprintf(f, ", ");
            }
            std::fprintf(f, "]");
            if (j < 3) std::fprintf(f, ",");
            std::fprintf(f, "\n");
        }
        std::fprintf(f, "    ]");
        if (i < 3) std::fprintf(f, ",");
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "  ]\n");
    std::fprintf(f, "}\n");
    std::fclose(f);
    return 0;
}

// ===========================================================================
// extern "C" surface for the Python tests (ctypes).  Complex arrays are
// interleaved (re, im) doubles.
// ===========================================================================
// ---------------------------------------------------------------------------
// Near-field singular corrections of the fast (FMM / pFFT) operators:
// BemFmmOperator::precompute_corrections, bem_fmm.cu:506-686 (FP64, as there).
// For every pair (m, n) of RWG functions with halves on one triangle, per
// wavenumber: C_L = ik sum f.f DG jw jw - (i/k) d d sum DG jw jw + the
// Graglia static part, with DG = -1/(4 pi R) at R > 1e-12 and the exact R = 0
// limit ik/(4 pi) (bem_fmm.cu:603-615); C_K = -sum gradG (f_m . (r_i - r_j) x
// f_n) jw jw over R > 1e-12 (bem_fmm.cu:664-680).  CSR rows sorted by column
// (bem_fmm.cu:540-567); values summed over the shared triangles in triangle
// order (the std::map of bem_fmm.cu:521-537).
// ---------------------------------------------------------------------------
struct NearCorr {
    std::vector<int> row_ptr, col;
    std::vector<cplx> L[2], K[2];  // [0] k_ext, [1] k_int
};

static NearCorr near_corrections(const Rwg& r, const MeshD& m, const Quad& q, cplx ke, cplx ki) {
    const int N = r.N, Nq = q.n;
    struct Hi { int n; double div, coeff; V3 fv; int sign; double area; };
    std::map<int, std::vector<Hi>> t2r;  // bem_fmm.cu:521-537
    for (int n = 0; n < N; n++) {
        t2r[r.tp[n]].push_back({n, r.len[n] / r.ap[n], r.len[n] / (2.0 * r.ap[n]), r.fp[n], +1, r.ap[n]});
        t2r[r.tm[n]].push_back({n, -r.len[n] / r.am[n], r.len[n] / (2.0 * r.am[n]), r.fm[n], -1, r.am[n]});
    }
    NearCorr C;
    std::vector<std::vector<int>> rows(N);  // bem_fmm.cu:540-555
    for (auto& kv : t2r)
        for (const Hi& a : kv.second)
            for (const Hi& b : kv.second) rows[a.n].push_back(b.n);
    C.row_ptr.assign(N + 1, 0);
    for (int i = 0; i < N; i++) {
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
        C.row_ptr[i + 1] = C.row_ptr[i] + (int)rows[i].size();
    }
    for (int i = 0; i < N; i++) C.col.insert(C.col.end(), rows[i].begin(), rows[i].end());
    const int nnz = C.row_ptr[N];
    for (int w = 0; w < 2; w++) {
        C.L[w].assign(nnz, cplx(0.0));
        C.K[w].assign(nnz, cplx(0.0));
    }
    const cplx I(0.0, 1.0);
    for (auto& kv : t2r) {  // bem_fmm.cu:573-686
        const std::vector<Hi>& hs = kv.second;
        V3 v0, v1, v2;
        m.tri(kv.first, v0, v1, v2);
        std::vector<V3> qp(Nq), V(Nq);
        std::vector<double> P(Nq), R((size_t)Nq * Nq);
        for (int i = 0; i < Nq; i++) {
            const double l0 = 1.0 - q.l1[i] - q.l2[i];
            qp[i] = v0 * l0 + v1 * q.l1[i] + v2 * q.l2[i];
        }
        for (int i = 0; i < Nq; i++)
            for (int j = 0; j < Nq; j++) R[(size_t)i * Nq + j] = (qp[i] - qp[j]).norm();
        for (int i = 0; i < Nq; i++) {
            P[i] = pot_integral(qp[i], v0, v1, v2);
            V[i] = vec_pot_integral(qp[i], v0, v1, v2, q);
        }
        // f and jw of each half on this triangle (bem_fmm.cu: f_p / f_m, jw_p / jw_m)
        std::vector<std::vector<V3>> f(hs.size(), std::vector<V3>(Nq));
        std::vector<std::vector<double>> jw(hs.size(), std::vector<double>(Nq));
        for (size_t h = 0; h < hs.size(); h++)
            for (int i = 0; i < Nq; i++) {
                f[h][i] = (qp[i] - hs[h].fv) * (hs[h].sign * hs[h].coeff);
                jw[h][i] = hs[h].area * q.w[i];
            }
        const cplx kv2[2] = {ke, ki};
        for (int w = 0; w < 2; w++) {
            const cplx k = kv2[w], ik = I * k, iok = I / k;
            std::vector<cplx> DG((size_t)Nq * Nq), gG((size_t)Nq * Nq, cplx(0.0));
            for (size_t e = 0; e < DG.size(); e++) {
                const double Rij = R[e];
                if (Rij > 1e-12) {
                    DG[e] = -1.0 / (4.0 * PI * Rij);
                    const cplx G = std::exp(ik * Rij) / (4.0 * PI * Rij);
                    gG[e] = G * (ik - 1.0 / Rij) / Rij;
                } else {
                    DG[e] = ik / (4.0 * PI);
                }
            }
            for (size_t a = 0; a < hs.size(); a++)
                for (size_t b = 0; b < hs.size(); b++) {
                    const int mrow = hs[a].n, ncol = hs[b].n;
                    const int* cb = C.col.data() + C.row_ptr[mrow];
                    const int* ce = C.col.data() + C.row_ptr[mrow + 1];
                    const int pos = C.row_ptr[mrow] + (int)(std::lower_bound(cb, ce, ncol) - cb);
                    cplx dv = 0.0, ds = 0.0;
                    for (int i = 0; i < Nq; i++)
                        for (int j = 0; j < Nq; j++) {
                            const double jp = jw[a][i] * jw[b][j];
                            const double fd = f[a][i].dot(f[b][j]);
                            dv += ik * fd * DG[(size_t)i * Nq + j] * jp;
                            ds += -iok * hs[a].div * hs[b].div * DG[(size_t)i * Nq + j] * jp;
                        }
                    cplx av = 0.0, as = 0.0;
                    for (int i = 0; i < Nq; i++) {
                        const V3 fo = (V[i] - hs[b].fv * P[i]) * (hs[b].sign * hs[b].coeff);
                        av += ik * f[a][i].dot(fo) * jw[a][i] * INV4PI;
                        as += -iok * hs[a].div * hs[b].div * P[i] * jw[a][i] * INV4PI;
                    }
                    C.L[w][pos] += dv + ds + av + as;
                    cplx kc = 0.0;
                    for (int i = 0; i < Nq; i++)
                        for (int j = 0; j < Nq; j++) {
                            if (R[(size_t)i * Nq + j] < 1e-12) continue;
                            const double jp = jw[a][i] * jw[b][j];
                            const V3 cr = (qp[i] - qp[j]).cross(f[b][j]);
                            kc += gG[(size_t)i * Nq + j] * f[a][i].dot(cr) * jp;
                        }
                    C.K[w][pos] -= kc;
                }
        }
    }
    return C;
}

struct orc_problem {
    MeshD mesh;
    Rwg rwg;
    Quad quad;
    double last_t_assemble = 0, last_t_solve = 0, last_t_farfield = 0;
};

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_set_r0_exact(int on) { g_r0_exact = on; }

int orc_quad_rule(int order, double* out /* 3*25: l1, l2, w */) {
    Quad q = quad_rule(order);
    for (int i = 0; i < q.n; i++) {
        out[i] = q.l1[i];
        out[25 + i] = q.l2[i];
        out[50 + i] = q.w[i];
    }
    return q.n;
}

// shape 0: icosphere(param = radius), 1: hex_prism(param = aspect ratio).
int orc_mesh_counts(int shape, int ref, int* nv, int* nt) {
    int T = (shape == 1 ? 24 : 20);
    for (int i = 0; i < ref; i++) T *= 4;
    *nt = T;
    *nv = T / 2 + 2;
    return 0;
}

int orc_make_mesh(int shape, int ref, double param, double* verts, int* tris) {
    MeshD m = shape == 1 ? hex_prism(param, ref) : icosphere(param, ref);
    for (size_t i = 0; i < m.v.size(); i++) {
        verts[3 * i] = m.v[i].x;
        verts[3 * i + 1] = m.v[i].y;
        verts[3 * i + 2] = m.v[i].z;
    }
    std::memcpy(tris, m.t.data(), m.t.size() * sizeof(int));
    return (int)m.v.size();
}

orc_problem* orc_create(const double* verts, int nv, const int* tris, int nt, int quad_order) {
    orc_problem* p = new orc_problem();
    p->mesh.v.resize(nv);
    for (int i = 0; i < nv; i++) p->mesh.v[i] = V3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    p->mesh.t.assign(tris, tris + 3 * nt);
    p->rwg = build_rwg(p->mesh);
    p->quad = quad_rule(quad_order);
    return p;
}

void orc_destroy(orc_problem* p) { delete p; }
int orc_N(const orc_problem* p) { return p->rwg.N; }
int orc_Nq(const orc_problem* p) { return p->quad.n; }

int orc_get_rwg(const orc_problem* p, int* tri_p, int* tri_m, double* free_p, double* free_m, double* len,
                double* area_p, double* area_m) {
    const Rwg& r = p->rwg;
    for (int n = 0; n < r.N; n++) {
        tri_p[n] = r.tp[n];
        tri_m[n] = r.tm[n];
        free_p[3 * n] = r.fp[n].x; free_p[3 * n + 1] = r.fp[n].y; free_p[3 * n + 2] = r.fp[n].z;
        free_m[3 * n] = r.fm[n].x; free_m[3 * n + 1] = r.fm[n].y; free_m[3 * n + 2] = r.fm[n].z;
        len[n] = r.len[n];
        area_p[n] = r.ap[n];
        area_m[n] = r.am[n];
    }
    return r.N;
}

double orc_potential_integral(const double* r, const double* a, const double* b, const double* c) {
    return pot_integral(V3(r[0], r[1], r[2]), V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]), V3(c[0], c[1], c[2]));
}

double orc_integral_R(const double* r, const double* a, const double* b, const double* c, int order) {
    return integral_R(V3(r[0], r[1], r[2]), V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]), V3(c[0], c[1], c[2]),
                      quad_rule(order));
}

void orc_vector_potential(const double* r, const double* a, const double* b, const double* c, int order,
                          double* out) {
    V3 v = vec_pot_integral(V3(r[0], r[1], r[2]), V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]),
                            V3(c[0], c[1], c[2]), quad_rule(order));
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

void orc_assemble_LK(orc_problem* p, double kre, double kim, double* L, double* K) {
    assemble_LK(p->rwg, p->mesh, p->quad, cplx(kre, kim), reinterpret_cast<cplx*>(L), reinterpret_cast<cplx*>(K));
}

void orc_assemble_pmchwt(orc_problem* p, double ker, double kei, double kir, double kii, double ee, double ei,
                         double* Z) {
    double t0 = now_s();
    assemble_pmchwt(p->rwg, p->mesh, p->quad, cplx(ker, kei), cplx(kir, kii), ee, ei, reinterpret_cast<cplx*>(Z));
    p->last_t_assemble = now_s() - t0;
}

void orc_rhs_planewave(orc_problem* p, double kre, double kim, double ee, const double* E0, const double* kh,
                       double* b) {
    rhs_planewave(p->rwg, p->mesh, p->quad, cplx(kre, kim), ee, V3(E0[0], E0[1], E0[2]), V3(kh[0], kh[1], kh[2]),
                  reinterpret_cast<cplx*>(b));
}

int orc_lu_solve(double* Z, int n, double* B, int nrhs) {
    std::vector<int> piv;
    int info = lu_factor_rowmajor(reinterpret_cast<cplx*>(Z), n, piv);
    if (info) return info;  // solver.cu:115-121 leaves B untouched on failure
    lu_solve_rowmajor(reinterpret_cast<cplx*>(Z), n, piv, reinterpret_cast<cplx*>(B), nrhs);
    return 0;
}

void orc_farfield_vec(orc_problem* p, const double* cJ, const double* cM, double kre, double kim, double ee,
                      const double* rh, int ndir, double* Fv) {
    std::vector<V3> r(ndir);
    for (int d = 0; d < ndir; d++) r[d] = V3(rh[3 * d], rh[3 * d + 1], rh[3 * d + 2]);
    farfield_vec(p->rwg, p->mesh, p->quad, reinterpret_cast<const cplx*>(cJ), reinterpret_cast<const cplx*>(cM),
                 cplx(kre, kim), ee, r.data(), ndir, reinterpret_cast<cplx*>(Fv));
}

void orc_amplitude_to_mueller(const double* S1, const double* S2, const double* S3, const double* S4, int nt,
                              double* M) {
    mueller(reinterpret_cast<const cplx*>(S1), reinterpret_cast<const cplx*>(S2), reinterpret_cast<const cplx*>(S3),
            reinterpret_cast<const cplx*>(S4), nt, M);
}

void orc_gauss_legendre(int n, double* x, double* w) {
    std::vector<double> xx, ww;
    gauss_legendre(n, xx, ww);
    for (int i = 0; i < n; i++) { x[i] = xx[i]; w[i] = ww[i]; }
}

int orc_orientations(int na, int nb, int ng, int bsym, int gsym, int mirror, int sort, double* RT, double* w) {
    std::vector<Orient> o = gen_orient(na, nb, ng, bsym, gsym, mirror != 0);
    if (sort) sort_orient(o);
    for (size_t i = 0; i < o.size(); i++) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) RT[9 * i + 3 * a + b] = o[i].RT.m[a][b];
        w[i] = o[i].w;
    }
    return (int)o.size();
}

// Single orientation: main.cpp:672-720.  X_out (nullable) receives the 2N x 2
// column-major solution.
int orc_dense_single(orc_problem* p, double ka, double mre, double mim, int nt, double* M, double* X_out) {
    const Rwg& r = p->rwg;
    int N = r.N, N2 = 2 * N;
    cplx m(mre, mim), ke(ka, 0.0), ki = ke * m;
    double ee = 1.0, ei = 1.0 / std::abs(m);
    double t0 = now_s();
    std::vector<cplx> Z((size_t)N2 * N2);
    assemble_pmchwt(r, p->mesh, p->quad, ke, ki, ee, ei, Z.data());
    double t1 = now_s();
    std::vector<cplx> B((size_t)N2 * 2);
    rhs_planewave(r, p->mesh, p->quad, ke, ee, V3(1, 0, 0), V3(0, 0, 1), B.data());
    rhs_planewave(r, p->mesh, p->quad, ke, ee, V3(0, 1, 0), V3(0, 0, 1), B.data() + N2);
    std::vector<int> piv;
    int info = lu_factor_rowmajor(Z.data(), N2, piv);
    if (!info) lu_solve_rowmajor(Z.data(), N2, piv, B.data(), 2);
    double t2 = now_s();
    if (X_out) std::memcpy(X_out, B.data(), sizeof(cplx) * N2 * 2);
    std::vector<V3> rh(nt);
    std::vector<double> th(nt);
    for (int t = 0; t < nt; t++) {
        th[t] = PI * t / (nt - 1);  // main.cpp:331-333
        rh[t] = V3(std::sin(th[t]), 0.0, std::cos(th[t]));
    }
    std::vector<cplx> Fp(nt * 3), Fq(nt * 3);
    farfield_vec(r, p->mesh, p->quad, B.data(), B.data() + N, ke, ee, rh.data(), nt, Fp.data());
    farfield_vec(r, p->mesh, p->quad, B.data() + N2, B.data() + N2 + N, ke, ee, rh.data(), nt, Fq.data());
    cplx ik = cplx(0, -1) * ke;
    std::vector<cplx> S1(nt), S2(nt), S3(nt), S4(nt);
    for (int t = 0; t < nt; t++) {  // farfield.cu:443-459 then main.cpp:703-711
        double ct = std::cos(th[t]), st = std::sin(th[t]);
        cplx fth_p = Fp[t * 3] * ct + Fp[t * 3 + 2] * (-st), fph_p = Fp[t * 3 + 1];
        cplx fth_q = Fq[t * 3] * ct + Fq[t * 3 + 2] * (-st), fph_q = Fq[t * 3 + 1];
        S2[t] = ik * fth_p;
        S4[t] = ik * fph_p;
        S3[t] = ik * fth_q;
        S1[t] = ik * fph_q;
    }
    mueller(S1.data(), S2.data(), S3.data(), S4.data(), nt, M);
    double k2 = std::norm(ke);
    for (int i = 0; i < 16 * nt; i++) M[i] /= k2;
    double t3 = now_s();
    p->last_t_assemble = t1 - t0;
    p->last_t_solve = t2 - t1;
    p->last_t_farfield = t3 - t2;
    return info;
}

// Orientation average: main.cpp:722-853.  Orientations generated, sorted, then
// restricted to [i0, i1) when 0 <= i0 < i1.
int orc_dense_orient(orc_problem* p, double ka, double mre, double mim, int na, int nb, int ng, int bsym, int gsym,
                     int mirror, int i0, int i1, int nt, double* M, double* X_out) {
    const Rwg& r = p->rwg;
    int N = r.N, N2 = 2 * N;
    cplx m(mre, mim), ke(ka, 0.0), ki = ke * m;
    double ee = 1.0, ei = 1.0 / std::abs(m);
    double t0 = now_s();
    std::vector<cplx> Z((size_t)N2 * N2);
    assemble_pmchwt(r, p->mesh, p->quad, ke, ki, ee, ei, Z.data());
    double t1 = now_s();
    std::vector<Orient> o = gen_orient(na, nb, ng, bsym, gsym, mirror != 0);
    sort_orient(o);
    int ntot = (int)o.size(), os = 0, oe = ntot;
    if (i0 >= 0 && i1 > i0) { os = std::min(i0, ntot); oe = std::min(i1, ntot); }
    int nr = oe - os;
    std::vector<cplx> B((size_t)N2 * nr * 2);
    for (int ri = 0; ri < nr; ri++) {
        const M3& RT = o[os + ri].RT;
        V3 kh = mul(RT, V3(0, 0, 1)), ep = mul(RT, V3(1, 0, 0)), eq = mul(RT, V3(0, 1, 0));
        rhs_planewave(r, p->mesh, p->quad, ke, ee, ep, kh, B.data() + (size_t)(2 * ri) * N2);
        rhs_planewave(r, p->mesh, p->quad, ke, ee, eq, kh, B.data() + (size_t)(2 * ri + 1) * N2);
    }
    std::vector<int> piv;
    int info = lu_factor_rowmajor(Z.data(), N2, piv);
    if (!info) lu_solve_rowmajor(Z.data(), N2, piv, B.data(), 2 * nr);
    double t2 = now_s();
    if (X_out) std::memcpy(X_out, B.data(), sizeof(cplx) * (size_t)N2 * 2 * nr);
    std::vector<double> th(nt);
    for (int t = 0; t < nt; t++) th[t] = PI * t / (nt - 1);
    for (int i = 0; i < 16 * nt; i++) M[i] = 0.0;
    cplx ik = cplx(0, -1) * ke;
    double k2 = std::norm(ke);
    std::vector<std::vector<double>> Mo(nr, std::vector<double>(16 * nt));
#pragma omp parallel for schedule(dynamic, 1)
    for (int ri = 0; ri < nr; ri++) {
        const M3& RT = o[os + ri].RT;
        std::vector<V3> rh(nt), ep(nt);
        V3 eq = mul(RT, V3(0, 1, 0));
        for (int t = 0; t < nt; t++) {
            double ct = std::cos(th[t]), st = std::sin(th[t]);
            rh[t] = mul(RT, V3(st, 0, ct));
            ep[t] = mul(RT, V3(ct, 0, -st));
        }
        std::vector<cplx> Fp(nt * 3), Fq(nt * 3);
        const cplx* Xp = B.data() + (size_t)(2 * ri) * N2;
        const cplx* Xq = B.data() + (size_t)(2 * ri + 1) * N2;
        farfield_vec(r, p->mesh, p->quad, Xp, Xp + N, ke, ee, rh.data(), nt, Fp.data());
        farfield_vec(r, p->mesh, p->quad, Xq, Xq + N, ke, ee, rh.data(), nt, Fq.data());
        std::vector<cplx> S1(nt), S2(nt), S3(nt), S4(nt);
        for (int t = 0; t < nt; t++) {  // main.cpp:817-831
            const V3& a = ep[t];
            cplx Fpp = Fp[t * 3] * a.x + Fp[t * 3 + 1] * a.y + Fp[t * 3 + 2] * a.z;
            cplx Fpq = Fp[t * 3] * eq.x + Fp[t * 3 + 1] * eq.y + Fp[t * 3 + 2] * eq.z;
            cplx Fqp = Fq[t * 3] * a.x + Fq[t * 3 + 1] * a.y + Fq[t * 3 + 2] * a.z;
            cplx Fqq = Fq[t * 3] * eq.x + Fq[t * 3 + 1] * eq.y + Fq[t * 3 + 2] * eq.z;
            S2[t] = ik * Fpp;
            S4[t] = ik * Fpq;
            S3[t] = ik * Fqp;
            S1[t] = ik * Fqq;
        }
        mueller(S1.data(), S2.data(), S3.data(), S4.data(), nt, Mo[ri].data());
        if (mirror) {  // main.cpp:837-842
            int odd[8] = {2, 3, 6, 7, 8, 9, 12, 13};
            for (int b = 0; b < 8; b++)
                for (int t = 0; t < nt; t++) Mo[ri][odd[b] * nt + t] = 0.0;
        }
    }
    for (int ri = 0; ri < nr; ri++) {  // main.cpp:844-845, summed in orientation order
        double w = o[os + ri].w;
        for (int i = 0; i < 16 * nt; i++) M[i] += w * Mo[ri][i] / k2;
    }
    double t3 = now_s();
    p->last_t_assemble = t1 - t0;
    p->last_t_solve = t2 - t1;
    p->last_t_farfield = t3 - t2;
    return info;
}

// Near-field corrections (bem_fmm.cu:506-686) for k_ext and k_int: returns
// nnz; with col_idx non-null also fills row_ptr (N + 1), col_idx (nnz) and the
// four value arrays (nnz interleaved complex each).
int orc_near_corrections(orc_problem* p, double ker, double kei, double kir, double kii, int* row_ptr, int* col_idx,
                         double* Lext, double* Kext, double* Lint, double* Kint) {
    NearCorr C = near_corrections(p->rwg, p->mesh, p->quad, cplx(ker, kei), cplx(kir, kii));
    const int nnz = C.row_ptr.back();
    if (row_ptr) std::memcpy(row_ptr, C.row_ptr.data(), C.row_ptr.size() * sizeof(int));
    if (col_idx) {
        std::memcpy(col_idx, C.col.data(), nnz * sizeof(int));
        double* out[4] = {Lext, Kext, Lint, Kint};
        const std::vector<cplx>* src[4] = {&C.L[0], &C.K[0], &C.L[1], &C.K[1]};
        for (int a = 0; a < 4; a++)
            if (out[a]) std::memcpy(out[a], src[a]->data(), nnz * sizeof(cplx));
    }
    return nnz;
}

void orc_last_timings(const orc_problem* p, double* out3) {
    out3[0] = p->last_t_assemble;
    out3[1] = p->last_t_solve;
    out3[2] = p->last_t_farfield;
}

int orc_write_json(const char* fn, const double* M, const double* th, int nt, double ka, double nre, double nim,
                   int ref, int na, int nb, int ng, double ta, double ts, double tf, double tt) {
    return write_json(fn, M, th, nt, ka, nre, nim, ref, na, nb, ng, ta, ts, tf, tt);
}

}  // extern "C"
