// abi.cpp -- implementation of include/bem_abi.h: context, mesh upload,
// device buffers, argument checking, host<->device layout conversion.
// The numerical work is in kernels/*.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <utility>

#include "bem_internal.h"

int bem_fail(bem_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int ctx_scratch(bem_ctx* c, size_t bytes, void** out) {
    if (c->scratch_bytes < bytes) {
        if (c->scratch) (void)hipFree(c->scratch);
        c->scratch = nullptr;
        c->scratch_bytes = 0;
        if (hipMalloc(&c->scratch, bytes) != hipSuccess) {
            c->scratch = nullptr;
            return bem_fail(c, BEM_ERR_NOMEM, "scratch: cannot allocate " + std::to_string(bytes) + " bytes");
        }
        c->scratch_bytes = bytes;
    }
    *out = c->scratch;
    return 0;
}

void kt_resolve(bem_ctx* c) {
    KernelTimer& k = c->kt;
    if (k.pending.empty()) return;
    (void)hipStreamSynchronize(c->stream3);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream);
    for (KernelTimer::Rec& r : k.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            k.ms[r.cls] += ms;
            k.flops[r.cls] += r.flops;
            k.bytes[r.cls] += r.bytes;
            k.launches[r.cls] += 1;
        }
        k.pool.push_back(r.a);
        k.pool.push_back(r.b);
    }
    k.pending.clear();
}

namespace {

// Resolves kernel timings on scope exit of an ABI call.
struct KtScope {
    bem_ctx* c;
    ~KtScope() {
        if (c) kt_resolve(c);
    }
};

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
int dalloc(bem_ctx* c, T*& p, size_t count, const char* what) {
    dfree(p);
    if (count == 0) return 0;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
        p = nullptr;
        return bem_fail(c, BEM_ERR_NOMEM,
                        std::string("cannot allocate ") + what + " (" + std::to_string(count * sizeof(T)) + " bytes)");
    }
    return 0;
}

void free_mesh(bem_ctx* c) {
    DevMesh& d = c->dm;
    dfree(d.qp);
    dfree(d.tv);
    dfree(d.hc);
    dfree(d.hp);
    dfree(d.hrwg);
    dfree(d.rwg_half);
    dfree(d.near);
    dfree(d.pv);
    d = DevMesh();
    bj_free(c);  // the blocks belong to the mesh
}

void free_matrices(bem_ctx* c) {
    dfree(c->Zr);
    dfree(c->Zi);
    dfree(c->Wr);
    dfree(c->Wi);
    dfree(c->Vr);
    dfree(c->Vi);
    dfree(c->ipiv);
    dfree(c->perm);
    if (c->inv32) (void)hipFree(c->inv32);
    c->inv32 = nullptr;
    c->inv32_bytes = 0;
    c->has_Z64 = c->has_Z32 = false;
    c->factored = -1;
    c->lu_avail = -1;
    c->factors_stale = false;
}

void free_rhs(bem_ctx* c) {
    dfree(c->Br);
    dfree(c->Bi);
    dfree(c->Xr);
    dfree(c->Xi);
    c->rhs_cap = 0;
    c->has_B = c->has_X = false;
}

bool finite3(const double* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// Sets the system size and (re)allocates the matrix buffers it needs.
int set_size(bem_ctx* c, int n) {
    const int npad = round_up(std::max(n, 1), 256);
    if (n != c->n || npad != c->npad) {
        free_matrices(c);
        free_rhs(c);
        c->n = n;
        c->npad = npad;
    }
    return 0;
}

int ensure_z64(bem_ctx* c) {
    const size_t nn = (size_t)c->npad * c->npad;
    if (!c->Zr) {
        if (int rc = dalloc(c, c->Zr, nn, "Z (FP64, real)")) return rc;
        if (int rc = dalloc(c, c->Zi, nn, "Z (FP64, imag)")) return rc;
    }
    return 0;
}

int ensure_w32(bem_ctx* c) {
    const size_t nn = (size_t)c->npad * c->npad;
    if (!c->Wr) {
        if (int rc = dalloc(c, c->Wr, nn, "Z (FP32, real)")) return rc;
        if (int rc = dalloc(c, c->Wi, nn, "Z (FP32, imag)")) return rc;
    }
    return 0;
}

struct EventTimer {
    bem_ctx* c;
    explicit EventTimer(bem_ctx* cc) : c(cc) { (void)hipEventRecord(c->ev[0], c->stream); }
    double stop() {
        (void)hipEventRecord(c->ev[1], c->stream);
        (void)hipEventSynchronize(c->ev[1]);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        return ms;
    }
};

}  // namespace

// The device solutions no longer belong to the device right-hand sides (new
// right-hand sides, or buffers reallocated or re-strided): orient_mueller and
// mueller_grad refuse to run until the next solve (and orient_mueller).
static void drop_solution(bem_ctx* c) {
    c->has_X = false;
    c->x_tol = 0.0;
    c->last_norient = 0;
}

// B/X buffers: planar FP64, row-major n x ldr (ldr multiple of 16), rows padded
// to npad.  Called before new right-hand sides are written: the old solutions
// are dropped.  On an allocation failure no buffer is left half-allocated.
int ctx_ensure_rhs(bem_ctx* c, int nrhs) {
    const int ldr = round_up(std::max(nrhs, 1), 16);
    const size_t need = (size_t)c->npad * ldr;
    drop_solution(c);
    if (need > c->rhs_cap) {
        c->has_B = false;
        c->rhs_cap = 0;
        int rc = dalloc(c, c->Br, need, "B (real)");
        if (!rc) rc = dalloc(c, c->Bi, need, "B (imag)");
        if (!rc) rc = dalloc(c, c->Xr, need, "X (real)");
        if (!rc) rc = dalloc(c, c->Xi, need, "X (imag)");
        if (rc) {
            dfree(c->Br);
            dfree(c->Bi);
            dfree(c->Xr);
            dfree(c->Xi);
            return rc;
        }
        c->rhs_cap = need;
    }
    if (ldr != c->ldr) c->has_B = false;  // the old right-hand sides have another stride
    c->nrhs = nrhs;
    c->ldr = ldr;
    return 0;
}

extern "C" {

bem_ctx* bem_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    bem_ctx* c = new bem_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[2], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[3], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[4], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->lu_ev[5], hipEventDisableTiming) != hipSuccess) {
        for (hipEvent_t e : {c->ev[0], c->ev[1]})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : c->lu_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t s : {c->stream, c->stream2, c->stream3})
            if (s) (void)hipStreamDestroy(s);
        delete c;
        return nullptr;
    }
    // the device address of the no-allocation array; a failure is reported by
    // bem_farfield_noalloc, the other entry points do not need it
    (void)ff_noalloc_init(c);
    c->err.clear();
    return c;
}

void bem_destroy(bem_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream3);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream);
    free_mesh(c);
    free_matrices(c);
    free_rhs(c);
    dfree(c->H);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->part) (void)hipFree(c->part);
    (void)hipEventDestroy(c->ev[0]);
    (void)hipEventDestroy(c->ev[1]);
    for (hipEvent_t e : c->lu_ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream3);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* bem_last_error(const bem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bem_timings(const bem_ctx* c, double* out, int n) {
    if (!c || !out || n < 0) return BEM_ERR_ARG;
    const int m = std::min(n, 8);
    for (int i = 0; i < m; i++) out[i] = c->t_ms[i];
    return m;
}

int bem_kernel_stats(bem_ctx* c, int mode, double* out, int n) {
    if (!c) return BEM_ERR_ARG;
    KernelTimer& k = c->kt;
    kt_resolve(c);
    if (mode == 0 || mode == 1 || mode == 3) {
        k.on = mode != 0;
        k.mask = mode == 3 ? ((1u << KS_UPDATE32) | (1u << KS_UPDATE64)) : ~0u;
        for (int i = 0; i < KS_COUNT; i++) {
            k.ms[i] = k.flops[i] = k.bytes[i] = 0.0;
            k.launches[i] = 0;
        }
        return KS_COUNT;
    }
    if (mode != 2 || !out) return BEM_ERR_ARG;
    const int m = std::min(n, (int)KS_COUNT);
    for (int i = 0; i < m; i++) {
        out[4 * i] = (double)k.launches[i];
        out[4 * i + 1] = k.ms[i];
        out[4 * i + 2] = k.flops[i];
        out[4 * i + 3] = k.bytes[i];
    }
    return m;
}

int bem_autotune(bem_ctx* c, int nrhs, double* out, int nout) {
    if (!c || nrhs < 1 || nout < 0) return bem_fail(c, BEM_ERR_ARG, "autotune: bad arguments");
    if (c->npad <= 0) return bem_fail(c, BEM_ERR_STATE, "autotune: no system size yet (set a mesh or a matrix)");
    (void)hipSetDevice(c->device);
    double local[32];
    return gemm_autotune_launch(c, c->npad, nrhs, out ? out : local, out ? nout : 32);
}

int bem_gemm_plan(bem_ctx* c, int prec, int large, int skinny, long cap) {
    if (!c) return BEM_ERR_ARG;
    (void)hipSetDevice(c->device);
    return gemm_set_plan(c, prec, large, skinny, cap);
}

int bem_gemm_plan_get(bem_ctx* c, int prec, double* out, int nout) {
    if (!c || prec < 0 || prec > 1 || !out || nout < 0) return bem_fail(c, BEM_ERR_ARG, "gemm_plan_get: bad arguments");
    const GemmPlan& p = c->gemm_plan[prec];
    double v[8 + GemmPlan::kMaxCfgs];
    int k = 0;
    v[k++] = p.large;
    v[k++] = p.skinny;
    v[k++] = p.panel;
    v[k++] = (double)p.cap;
    v[k++] = p.tuned ? 1.0 : 0.0;
    v[k++] = p.trialed ? 1.0 : 0.0;
    for (double t : p.trial_ms) v[k++] = t;
    v[k++] = prec == 0 ? p.nt : p.large;  // the trailing update's configuration
    const int m = std::min(k, nout);
    for (int i = 0; i < m; i++) out[i] = v[i];
    return m;
}

int bem_device_copy(bem_ctx* c, double bytes) {
    if (!c || !(bytes >= 16.0) || bytes > 64e9) return bem_fail(c, BEM_ERR_ARG, "device_copy: bad size");
    (void)hipSetDevice(c->device);
    return device_copy_launch(c, (size_t)bytes / 16);
}

int bem_mesh_counts(int shape, int refinements, int* nv, int* nt) {
    if (!nv || !nt || refinements < 0 || refinements > 8 || (shape != 0 && shape != 1)) return BEM_ERR_ARG;
    long t = shape == 1 ? 24 : 20;
    for (int i = 0; i < refinements; i++) t *= 4;
    *nt = (int)t;
    *nv = (int)(t / 2 + 2);  // closed genus-0 surface: V = F/2 + 2
    return 0;
}

int bem_make_mesh(int shape, int refinements, double param, double* verts, int* tris) {
    int nv = 0, nt = 0;
    if (!verts || !tris || bem_mesh_counts(shape, refinements, &nv, &nt) || !(param > 0.0)) return BEM_ERR_ARG;
    const Mesh m = shape == 1 ? hex_prism(param, refinements) : icosphere(param, refinements);
    if (m.nv() != nv || m.nt() != nt) return BEM_ERR_MESH;
    for (int i = 0; i < nv; i++) {
        verts[3 * i] = m.verts[i].x;
        verts[3 * i + 1] = m.verts[i].y;
        verts[3 * i + 2] = m.verts[i].z;
    }
    std::memcpy(tris, m.tris.data(), sizeof(int) * 3 * (size_t)nt);
    return 0;
}

int bem_write_json(const char* path, const double* M, const double* theta, int ntheta, double ka, double n_re,
                   double n_im, int refinements, int n_alpha, int n_beta, int n_gamma, const double* timing) {
    if (!path || !M || !theta || ntheta < 1 || !timing) return BEM_ERR_ARG;
    return write_json(path, M, theta, ntheta, ka, n_re, n_im, refinements, n_alpha, n_beta, n_gamma, timing[0],
                      timing[1], timing[2], timing[3])
               ? 0
               : BEM_ERR_ARG;
}

int bem_write_per_orient(const char* path, int n_orient, int ntheta, int n_alpha, int n_beta, const double* weights,
                         const double* M_orient) {
    if (!path || n_orient < 0 || ntheta < 1 || (n_orient > 0 && (!weights || !M_orient))) return BEM_ERR_ARG;
    return write_per_orient(path, n_orient, ntheta, n_alpha, n_beta, weights, M_orient) ? 0 : BEM_ERR_ARG;
}

int bem_orientations(int na, int nb, int ng, int bsym, int gsym, int mirror, int sort, double* RT, double* w) {
    if (na < 1 || nb < 1 || ng < 1 || (bsym != 1 && bsym != 2) || gsym < 1 || !RT || !w) return BEM_ERR_ARG;
    std::vector<Orientation> o = generate_orientations(na, nb, ng, bsym, gsym, mirror != 0);
    if (sort) sort_orientations_nearest(o);
    for (size_t i = 0; i < o.size(); i++) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) RT[9 * i + 3 * a + b] = o[i].RT.m[a][b];
        w[i] = o[i].weight;
    }
    return (int)o.size();
}

int bem_set_mesh(bem_ctx* c, const double* verts, int nv, const int* tris, int nt, int quad_order) {
    if (!c) return BEM_ERR_ARG;
    if (!verts || !tris || nv < 3 || nt < 1) return bem_fail(c, BEM_ERR_ARG, "set_mesh: empty mesh");
    for (int i = 0; i < nv; i++)
        if (!finite3(verts + 3 * i)) return bem_fail(c, BEM_ERR_MESH, "set_mesh: non-finite vertex");
    for (int i = 0; i < 3 * nt; i++)
        if (tris[i] < 0 || tris[i] >= nv) return bem_fail(c, BEM_ERR_ARG, "set_mesh: vertex index out of range");
    (void)hipSetDevice(c->device);
    Mesh m;
    m.verts.resize(nv);
    for (int i = 0; i < nv; i++) m.verts[i] = Vec3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.tris.assign(tris, tris + 3 * nt);
    RWG r = build_rwg(m);
    if (r.N == 0) return bem_fail(c, BEM_ERR_MESH, "set_mesh: no interior edge (no RWG basis function)");
    free_mesh(c);
    c->mesh = m;
    c->rwg = r;
    c->quad = tri_quadrature(quad_order);
    c->has_mesh = true;
    c->asm_valid = false;
    const int T = nt, Nq = c->quad.npts, N = r.N;
    DevMesh& d = c->dm;
    d.T = T;
    d.N = N;
    d.Nq = Nq;

    // quadrature points, same arithmetic as precompute_half (assembly.cu:180-197)
    std::vector<double4> qp((size_t)T * Nq);
    std::vector<int4> tv(T);
    for (int t = 0; t < T; t++) {
        Vec3 v0, v1, v2;
        m.tri_verts(t, v0, v1, v2);
        const double A = m.tri_area(t);
        for (int q = 0; q < Nq; q++) {
            const double l1 = c->quad.pts[q][0], l2 = c->quad.pts[q][1], l0 = 1.0 - l1 - l2;
            const Vec3 rr = v0 * l0 + v1 * l1 + v2 * l2;
            qp[(size_t)t * Nq + q] = make_double4(rr.x, rr.y, rr.z, A * c->quad.wts[q]);
        }
        tv[t] = make_int4(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], 0);
    }
    // halves: slot j of triangle t is edge (j, j+1) with free vertex j+2
    std::map<std::pair<int, int>, int> edge_rwg;
    {
        int idx = 0;
        std::map<std::pair<int, int>, int> count;
        for (int t = 0; t < T; t++)
            for (int j = 0; j < 3; j++) {
                const int a = tris[3 * t + j], b = tris[3 * t + (j + 1) % 3];
                count[{std::min(a, b), std::max(a, b)}]++;
            }
        for (const auto& e : count)
            if (e.second == 2) edge_rwg[e.first] = idx++;
        if (idx != N) return bem_fail(c, BEM_ERR_MESH, "set_mesh: inconsistent edge map");
    }
    std::vector<double> hc((size_t)T * 3, 0.0), hp((size_t)T * 9, 0.0);
    std::vector<int> hr((size_t)T * 3, -1);
    std::vector<int4> rh(N, make_int4(-1, -1, -1, -1));
    for (int t = 0; t < T; t++)
        for (int j = 0; j < 3; j++) {
            const int a = tris[3 * t + j], b = tris[3 * t + (j + 1) % 3], f = tris[3 * t + (j + 2) % 3];
            auto it = edge_rwg.find({std::min(a, b), std::max(a, b)});
            if (it == edge_rwg.end()) continue;
            const int n = it->second;
            hr[t * 3 + j] = n;
            const bool plus = r.tri_p[n] == t;
            const double A = plus ? r.area_p[n] : r.area_m[n];
            hc[t * 3 + j] = (plus ? 1.0 : -1.0) * r.length[n] / (2.0 * A);
            hp[(t * 3 + j) * 3] = m.verts[f].x;
            hp[(t * 3 + j) * 3 + 1] = m.verts[f].y;
            hp[(t * 3 + j) * 3 + 2] = m.verts[f].z;
            if (plus) {
                rh[n].x = t;
                rh[n].y = j;
            } else {
                rh[n].z = t;
                rh[n].w = j;
            }
        }
    // triangle pairs sharing a vertex (incl. self), t <= s
    std::vector<std::vector<int>> vt(nv);
    for (int t = 0; t < T; t++)
        for (int j = 0; j < 3; j++) vt[tris[3 * t + j]].push_back(t);
    std::vector<int2> near;
    for (int t = 0; t < T; t++) {
        std::set<int> s;
        for (int j = 0; j < 3; j++)
            for (int u : vt[tris[3 * t + j]])
                if (u >= t) s.insert(u);
        for (int u : s) near.push_back(make_int2(t, u));
    }
    d.n_near = (int)near.size();

    int rc = 0;
    if ((rc = dalloc(c, d.qp, qp.size(), "quadrature points")) || (rc = dalloc(c, d.tv, tv.size(), "triangles")) ||
        (rc = dalloc(c, d.hc, hc.size(), "half coefficients")) || (rc = dalloc(c, d.hp, hp.size(), "free vertices")) ||
        (rc = dalloc(c, d.hrwg, hr.size(), "half RWG ids")) || (rc = dalloc(c, d.rwg_half, rh.size(), "RWG halves")) || (rc = dalloc(c, d.near, near.size(), "near pairs")) ||
        (rc = dalloc(c, d.pv, qp.size(), "Graglia values")))
        return rc;
    BEM_HIP(c, hipMemcpy(d.qp, qp.data(), qp.size() * sizeof(double4), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.tv, tv.data(), tv.size() * sizeof(int4), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.hc, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.hp, hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.hrwg, hr.data(), hr.size() * sizeof(int), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.rwg_half, rh.data(), rh.size() * sizeof(int4), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(d.near, near.data(), near.size() * sizeof(int2), hipMemcpyHostToDevice));
    return set_size(c, 2 * N);
}

int bem_num_rwg(const bem_ctx* c) { return (c && c->has_mesh) ? c->rwg.N : 0; }

int bem_get_rwg(const bem_ctx* c, int* tri_p, int* tri_m, double* free_p, double* free_m, double* length,
                double* area_p, double* area_m) {
    if (!c || !c->has_mesh) return BEM_ERR_STATE;
    const RWG& r = c->rwg;
    for (int n = 0; n < r.N; n++) {
        if (tri_p) tri_p[n] = r.tri_p[n];
        if (tri_m) tri_m[n] = r.tri_m[n];
        if (free_p) { free_p[3 * n] = r.free_p[n].x; free_p[3 * n + 1] = r.free_p[n].y; free_p[3 * n + 2] = r.free_p[n].z; }
        if (free_m) { free_m[3 * n] = r.free_m[n].x; free_m[3 * n + 1] = r.free_m[n].y; free_m[3 * n + 2] = r.free_m[n].z; }
        if (length) length[n] = r.length[n];
        if (area_p) area_p[n] = r.area_p[n];
        if (area_m) area_m[n] = r.area_m[n];
    }
    return r.N;
}

int bem_assemble(bem_ctx* c, double ker, double kei, double kir, double kii, double ee, double ei, unsigned flags) {
    if (!c) return BEM_ERR_ARG;
    KtScope kts{c};
    if (!c->has_mesh) return bem_fail(c, BEM_ERR_STATE, "assemble: no mesh");
    if (!(std::isfinite(ker) && std::isfinite(kei) && std::isfinite(kir) && std::isfinite(kii)) ||
        std::hypot(ker, kei) == 0.0 || std::hypot(kir, kii) == 0.0)
        return bem_fail(c, BEM_ERR_ARG, "assemble: wavenumbers must be finite and non-zero");
    if (!(ee > 0.0) || !(ei > 0.0) || !std::isfinite(ee) || !std::isfinite(ei))
        return bem_fail(c, BEM_ERR_ARG, "assemble: impedances must be positive");
    if (kei < 0.0 || kii < 0.0) return bem_fail(c, BEM_ERR_ARG, "assemble: Im k must be >= 0 (lossy media)");
    (void)hipSetDevice(c->device);
    if (flags == 0) flags = BEM_ASM_Z64 | BEM_ASM_Z32;
    if (int rc = set_size(c, 2 * c->rwg.N)) return rc;
    const bool w64 = flags & BEM_ASM_Z64, w32 = flags & BEM_ASM_Z32;
    if (w64)
        if (int rc = ensure_z64(c)) return rc;
    if (w32)
        if (int rc = ensure_w32(c)) return rc;
    AsmParams p;
    p.kr[0] = ker; p.ki[0] = kei; p.kr[1] = kir; p.ki[1] = kii;
    p.eta[0] = ee; p.eta[1] = ei;
    EventTimer tm(c);
    if (int rc = asm_launch(c, p, w64, w32)) return rc;
    c->t_ms[0] = tm.stop();
    BEM_HIP(c, hipGetLastError());
    c->has_Z64 = w64;
    c->has_Z32 = w32;
    // FP32 factors live in the FP32 working copy, which a Z32 assembly
    // overwrites; other factors stay as the preconditioner of
    // bem_solve_gmres but belong to an earlier matrix
    if (w32 && c->lu_avail == BEM_F32_3M) c->lu_avail = -1;
    if (c->factored == BEM_F32_3M && w32)
        c->factored = -1;
    else if (c->factored >= 0)
        c->factors_stale = true;
    c->asm_params = p;
    c->asm_valid = true;
    return 0;
}

int bem_get_Z(bem_ctx* c, double* Z, int row_major) {
    if (!c || !Z) return BEM_ERR_ARG;
    if (!c->has_Z64) return bem_fail(c, BEM_ERR_STATE, "get_Z: no FP64 Z on the device");
    const int n = c->n, ld = c->npad;
    std::vector<double> re((size_t)n * ld), im((size_t)n * ld);
    BEM_HIP(c, hipMemcpy(re.data(), c->Zr, re.size() * sizeof(double), hipMemcpyDeviceToHost));
    BEM_HIP(c, hipMemcpy(im.data(), c->Zi, im.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const size_t o = row_major ? (size_t)i * n + j : (size_t)j * n + i;
            Z[2 * o] = re[(size_t)i * ld + j];
            Z[2 * o + 1] = im[(size_t)i * ld + j];
        }
    return 0;
}

int bem_get_Z_entries(bem_ctx* c, const int* rows, const int* cols, int count, double* out) {
    if (!c || !rows || !cols || !out || count < 0) return bem_fail(c, BEM_ERR_ARG, "get_Z_entries: bad arguments");
    if (!c->has_Z64) return bem_fail(c, BEM_ERR_STATE, "get_Z_entries: no FP64 Z on the device");
    for (int e = 0; e < count; e++)
        if (rows[e] < 0 || rows[e] >= c->n || cols[e] < 0 || cols[e] >= c->n)
            return bem_fail(c, BEM_ERR_ARG, "get_Z_entries: index out of range");
    if (count == 0) return 0;
    (void)hipSetDevice(c->device);
    const size_t ib = ((size_t)count * sizeof(int) + 255) / 256 * 256;
    void* s = nullptr;
    if (int rc = ctx_scratch(c, 2 * ib + 2 * (size_t)count * sizeof(double), &s)) return rc;
    int* dr = static_cast<int*>(s);
    int* dc = reinterpret_cast<int*>(static_cast<char*>(s) + ib);
    double* dout = reinterpret_cast<double*>(static_cast<char*>(s) + 2 * ib);
    BEM_HIP(c, hipMemcpyAsync(dr, rows, (size_t)count * sizeof(int), hipMemcpyHostToDevice, c->stream));
    BEM_HIP(c, hipMemcpyAsync(dc, cols, (size_t)count * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (int rc = z_entries_launch(c, dr, dc, count, dout)) return rc;
    BEM_HIP(c, hipMemcpyAsync(out, dout, 2 * (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int bem_near_corrections(bem_ctx* c, double ker, double kei, double kir, double kii, int* row_ptr, int* col_idx,
                         double* L_ext, double* K_ext, double* L_int, double* K_int) {
    if (!c || !row_ptr) return bem_fail(c, BEM_ERR_ARG, "near_corrections: bad arguments");
    if (!c->has_mesh) return bem_fail(c, BEM_ERR_STATE, "near_corrections: no mesh");
    if ((ker == 0.0 && kei == 0.0) || (kir == 0.0 && kii == 0.0))
        return bem_fail(c, BEM_ERR_ARG, "near_corrections: zero wavenumber");
    (void)hipSetDevice(c->device);
    double* const out[4] = {L_ext, K_ext, L_int, K_int};
    return near_corr_launch(c, ker, kei, kir, kii, row_ptr, col_idx, out);
}

int bem_assemble_LK(bem_ctx* c, double kr, double ki, double* L, double* K) {
    if (!c || !L || !K) return BEM_ERR_ARG;
    if (int rc = bem_assemble(c, kr, ki, kr, ki, 1.0, 1.0, BEM_ASM_Z64)) return rc;
    // with equal wavenumbers and eta = 1: Z_TL = 2 L, Z_BL = 2 K
    const int N = c->rwg.N, ld = c->npad;
    std::vector<double> re((size_t)c->n * ld), im((size_t)c->n * ld);
    BEM_HIP(c, hipMemcpy(re.data(), c->Zr, re.size() * sizeof(double), hipMemcpyDeviceToHost));
    BEM_HIP(c, hipMemcpy(im.data(), c->Zi, im.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const size_t o = (size_t)i * N + j;
            L[2 * o] = 0.5 * re[(size_t)i * ld + j];
            L[2 * o + 1] = 0.5 * im[(size_t)i * ld + j];
            K[2 * o] = 0.5 * re[(size_t)(N + i) * ld + j];
            K[2 * o + 1] = 0.5 * im[(size_t)(N + i) * ld + j];
        }
    c->has_Z64 = false;  // the matrix on the device is not a PMCHWT system
    c->asm_valid = false;
    return 0;
}

int bem_set_Z(bem_ctx* c, const double* Z, int n, int row_major) {
    if (!c || !Z || n < 1) return bem_fail(c, BEM_ERR_ARG, "set_Z: bad arguments");
    (void)hipSetDevice(c->device);
    if (int rc = set_size(c, n)) return rc;
    if (int rc = ensure_z64(c)) return rc;
    const int ld = c->npad;
    std::vector<double> re((size_t)ld * ld, 0.0), im((size_t)ld * ld, 0.0);
    for (int i = 0; i < ld; i++)
        for (int j = 0; j < ld; j++) {
            if (i < n && j < n) {
                const size_t o = row_major ? (size_t)i * n + j : (size_t)j * n + i;
                re[(size_t)i * ld + j] = Z[2 * o];
                im[(size_t)i * ld + j] = Z[2 * o + 1];
            } else if (i == j) {
                re[(size_t)i * ld + j] = 1.0;
            }
        }
    BEM_HIP(c, hipMemcpy(c->Zr, re.data(), re.size() * sizeof(double), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(c->Zi, im.data(), im.size() * sizeof(double), hipMemcpyHostToDevice));
    c->has_Z64 = true;
    c->has_Z32 = false;
    if (c->factored >= 0) c->factors_stale = true;  // kept for bem_solve_gmres
    c->asm_valid = false;
    return 0;
}

int bem_factor(bem_ctx* c, int precision) {
    if (!c) return BEM_ERR_ARG;
    KtScope kts{c};
    if (precision != BEM_F64 && precision != BEM_F32_3M) return bem_fail(c, BEM_ERR_ARG, "factor: bad precision");
    if (!c->has_Z64 && !(precision == BEM_F32_3M && c->has_Z32))
        return bem_fail(c, BEM_ERR_STATE, "factor: no matrix (assemble or set_Z first)");
    (void)hipSetDevice(c->device);
    const size_t nn = (size_t)c->npad * c->npad;
    if (precision == BEM_F64 && !c->Vr) {
        if (int rc = dalloc(c, c->Vr, nn, "LU (FP64, real)")) return rc;
        if (int rc = dalloc(c, c->Vi, nn, "LU (FP64, imag)")) return rc;
    }
    if (precision == BEM_F32_3M)
        if (int rc = ensure_w32(c)) return rc;
    if (!c->ipiv) {
        if (int rc = dalloc(c, c->ipiv, c->npad, "pivots")) return rc;
        if (int rc = dalloc(c, c->perm, c->npad, "permutation")) return rc;
    }
    EventTimer tm(c);
    const int info = lu_factor_launch(c, precision);
    c->t_ms[1] = tm.stop();
    if (info < 0) return info;
    // the FP32 working copy now holds factors
    if (precision == BEM_F32_3M) c->has_Z32 = false;
    c->factored = info == 0 ? precision : -1;
    c->lu_avail = precision;  // also after a zero pivot (Zgetrf's factors, kept for get_lu)
    c->factors_stale = false;
    return info;
}

int bem_get_pivots(const bem_ctx* c, int* ipiv) {
    if (!c || !ipiv) return BEM_ERR_ARG;
    if (c->lu_avail < 0) return BEM_ERR_STATE;
    std::vector<int> h(c->npad);
    if (hipMemcpy(h.data(), c->ipiv, c->npad * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return BEM_ERR_HIP;
    for (int i = 0; i < c->n; i++) ipiv[i] = h[i] + 1;
    return 0;
}

int bem_get_lu(bem_ctx* c, double* LU) {
    if (!c || !LU) return BEM_ERR_ARG;
    if (c->lu_avail != BEM_F64) return bem_fail(c, BEM_ERR_STATE, "get_lu: no FP64 factorisation");
    const int n = c->n, ld = c->npad;
    std::vector<double> re((size_t)n * ld), im((size_t)n * ld);
    BEM_HIP(c, hipMemcpy(re.data(), c->Vr, re.size() * sizeof(double), hipMemcpyDeviceToHost));
    BEM_HIP(c, hipMemcpy(im.data(), c->Vi, im.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            LU[2 * ((size_t)i * n + j)] = re[(size_t)i * ld + j];
            LU[2 * ((size_t)i * n + j) + 1] = im[(size_t)i * ld + j];
        }
    return 0;
}

int bem_set_lu(bem_ctx* c, const double* LU, const int* ipiv, int n) {
    if (!c || !LU || !ipiv || n < 1) return bem_fail(c, BEM_ERR_ARG, "set_lu: bad arguments");
    for (int i = 0; i < n; i++)
        if (ipiv[i] < i + 1 || ipiv[i] > n) return bem_fail(c, BEM_ERR_ARG, "set_lu: pivot out of range");
    (void)hipSetDevice(c->device);
    if (int rc = set_size(c, n)) return rc;
    const int ld = c->npad;
    const size_t nn = (size_t)ld * ld;
    if (!c->Vr) {
        if (int rc = dalloc(c, c->Vr, nn, "LU (FP64, real)")) return rc;
        if (int rc = dalloc(c, c->Vi, nn, "LU (FP64, imag)")) return rc;
    }
    if (!c->ipiv) {
        if (int rc = dalloc(c, c->ipiv, ld, "pivots")) return rc;
        if (int rc = dalloc(c, c->perm, ld, "permutation")) return rc;
    }
    std::vector<double> re(nn, 0.0), im(nn, 0.0);
    for (int i = 0; i < ld; i++)
        for (int j = 0; j < ld; j++) {
            if (i < n && j < n) {
                re[(size_t)i * ld + j] = LU[2 * ((size_t)i * n + j)];
                im[(size_t)i * ld + j] = LU[2 * ((size_t)i * n + j) + 1];
            } else if (i == j) {
                re[(size_t)i * ld + j] = 1.0;
            }
        }
    std::vector<int> ip(ld), perm(ld);
    for (int i = 0; i < ld; i++) ip[i] = i < n ? ipiv[i] - 1 : i;
    for (int i = 0; i < ld; i++) perm[i] = i;
    for (int i = 0; i < ld; i++)
        if (ip[i] != i) std::swap(perm[i], perm[ip[i]]);
    BEM_HIP(c, hipMemcpy(c->Vr, re.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(c->Vi, im.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(c->ipiv, ip.data(), ld * sizeof(int), hipMemcpyHostToDevice));
    BEM_HIP(c, hipMemcpy(c->perm, perm.data(), ld * sizeof(int), hipMemcpyHostToDevice));
    c->factored = BEM_F64;
    c->lu_avail = BEM_F64;
    c->factors_stale = false;
    c->has_Z64 = false;  // no system matrix: solves run without refinement
    return 0;
}

// Right-hand sides from the host (column-major n x nrhs complex) into the
// device planes, or the device's own (from bem_rhs_planewave) when B is null.
static int load_rhs(bem_ctx* c, const double* B, int& nrhs, const char* who) {
    const int n = c->n;
    if (B) {
        if (nrhs < 1) return bem_fail(c, BEM_ERR_ARG, std::string(who) + ": nrhs must be positive");
        if (int rc = ctx_ensure_rhs(c, nrhs)) return rc;
        const int ld = c->ldr;
        std::vector<double> re((size_t)c->npad * ld, 0.0), im((size_t)c->npad * ld, 0.0);
        for (int j = 0; j < nrhs; j++)
            for (int i = 0; i < n; i++) {
                re[(size_t)i * ld + j] = B[2 * ((size_t)j * n + i)];
                im[(size_t)i * ld + j] = B[2 * ((size_t)j * n + i) + 1];
            }
        BEM_HIP(c, hipMemcpy(c->Br, re.data(), re.size() * sizeof(double), hipMemcpyHostToDevice));
        BEM_HIP(c, hipMemcpy(c->Bi, im.data(), im.size() * sizeof(double), hipMemcpyHostToDevice));
        c->has_B = true;
        c->rhs_planewave = false;
    } else {
        if (!c->has_B) return bem_fail(c, BEM_ERR_STATE, std::string(who) + ": no right-hand sides on the device");
        nrhs = c->nrhs;
    }
    return 0;
}

// The device solutions into X (column-major n x nrhs complex).
static int store_solution(bem_ctx* c, int nrhs, double* X) {
    const int n = c->n, ld = c->ldr;
    std::vector<double> re((size_t)n * ld), im((size_t)n * ld);
    BEM_HIP(c, hipMemcpy(re.data(), c->Xr, re.size() * sizeof(double), hipMemcpyDeviceToHost));
    BEM_HIP(c, hipMemcpy(im.data(), c->Xi, im.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < nrhs; j++)
        for (int i = 0; i < n; i++) {
            X[2 * ((size_t)j * n + i)] = re[(size_t)i * ld + j];
            X[2 * ((size_t)j * n + i) + 1] = im[(size_t)i * ld + j];
        }
    return 0;
}

int bem_solve(bem_ctx* c, const double* B, int nrhs, double* X, double ir_tol, int ir_max, bem_solve_stats* st) {
    if (!c) return BEM_ERR_ARG;
    KtScope kts{c};
    if (c->factored < 0) return bem_fail(c, BEM_ERR_STATE, "solve: no factorisation");
    if (c->factors_stale)
        return bem_fail(c, BEM_ERR_STATE, "solve: the factors belong to an earlier matrix (factor again, or use "
                                          "bem_solve_gmres)");
    if (ir_max < 0 || !(ir_tol >= 0.0)) return bem_fail(c, BEM_ERR_ARG, "solve: bad refinement parameters");
    (void)hipSetDevice(c->device);
    if (int rc = load_rhs(c, B, nrhs, "solve")) return rc;
    int iters = 0;
    double r0 = 0, r1 = 0, ts = 0, ti = 0;
    const int rc = lu_solve_launch(c, nrhs, ir_tol, ir_max, &iters, &r0, &r1, &ts, &ti);
    if (rc < 0 && rc != BEM_ERR_NOCONV) return rc;
    c->t_ms[2] = ts;
    c->t_ms[3] = ti;
    c->has_X = true;
    c->x_tol = ir_max > 0 ? ir_tol : 0.0;
    if (st) {
        st->iterations = iters;
        st->residual0 = r0;
        st->residual = r1;
        st->converged = (ir_max == 0 || r1 <= ir_tol) ? 1 : 0;
        st->t_solve_ms = ts;
        st->t_ir_ms = ti;
    }
    if (X)
        if (int e = store_solution(c, nrhs, X)) return e;
    return rc == BEM_ERR_NOCONV ? bem_fail(c, rc, "solve: refinement did not reach the tolerance") : 0;
}

int bem_solve_gmres(bem_ctx* c, const double* B, int nrhs, double* X, double tol, int restart, int max_iter,
                    int precond, bem_solve_stats* st) {
    if (!c) return BEM_ERR_ARG;
    KtScope kts{c};
    if (precond != BEM_PRECOND_LU && precond != BEM_PRECOND_BLOCK && precond != BEM_PRECOND_NONE)
        return bem_fail(c, BEM_ERR_ARG, "solve_gmres: unknown preconditioner");
    if (precond == BEM_PRECOND_LU && c->factored < 0)
        return bem_fail(c, BEM_ERR_STATE, "solve_gmres: no factors for the preconditioner");
    if (precond == BEM_PRECOND_BLOCK && (!c->bj.valid || 2 * c->bj.N != c->n))
        return bem_fail(c, BEM_ERR_STATE, "solve_gmres: no block-Jacobi preconditioner for this system");
    if (!c->has_Z64) return bem_fail(c, BEM_ERR_STATE, "solve_gmres: no FP64 matrix on the device");
    if (!(tol > 0.0) || restart < 1 || restart > 200 || max_iter < 1)
        return bem_fail(c, BEM_ERR_ARG, "solve_gmres: bad tolerance, restart length or iteration limit");
    (void)hipSetDevice(c->device);
    if (int rc = load_rhs(c, B, nrhs, "solve_gmres")) return rc;
    int iters = 0;
    double r0 = 0, r1 = 0, tms = 0;
    const int rc = gmres_solve_launch(c, nrhs, tol, restart, max_iter, precond, &iters, &r0, &r1, &tms);
    if (rc < 0 && rc != BEM_ERR_NOCONV) return rc;
    c->t_ms[2] = tms;
    c->t_ms[3] = 0.0;
    c->has_X = true;
    c->x_tol = tol;
    if (st) {
        st->iterations = iters;
        st->residual0 = r0;
        st->residual = r1;
        st->converged = rc == 0 ? 1 : 0;
        st->t_solve_ms = tms;
        st->t_ir_ms = 0.0;
    }
    if (X)
        if (int e = store_solution(c, nrhs, X)) return e;
    return rc == BEM_ERR_NOCONV ? bem_fail(c, rc, "solve_gmres: no convergence within the iteration limit") : 0;
}

int bem_block_jacobi(bem_ctx* c, double radius_mult, int max_block_rwg, int overlap_layers, double* build_ms) {
    if (!c) return BEM_ERR_ARG;
    if (!(radius_mult > 0.0) || max_block_rwg < 1 || overlap_layers < 0)
        return bem_fail(c, BEM_ERR_ARG, "block_jacobi: bad parameters");
    if (!c->has_mesh || !c->has_Z64 || c->n != 2 * c->rwg.N)
        return bem_fail(c, BEM_ERR_STATE, "block_jacobi: needs the mesh and its FP64 Z on the device");
    (void)hipSetDevice(c->device);
    double t = 0.0;
    const int nb = bj_build_launch(c, radius_mult, max_block_rwg, overlap_layers, &t);
    if (nb < 0) return nb;
    if (build_ms) *build_ms = t;
    return nb;
}

int bem_block_partition(const double* verts, int nv, const int* tris, int nt, int quad_order, double radius_mult,
                        int max_block_rwg, int overlap_layers, int* n_ext, int* sizes, int* rwg_ext) {
    if (!verts || !tris || nv < 3 || nt < 1 || !n_ext || !(radius_mult > 0.0) || max_block_rwg < 1 ||
        overlap_layers < 0)
        return BEM_ERR_ARG;
    Mesh m;
    m.verts.resize(nv);
    for (int i = 0; i < nv; i++) m.verts[i] = Vec3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.tris.assign(tris, tris + 3 * (size_t)nt);
    for (int v : m.tris)
        if (v < 0 || v >= nv) return BEM_ERR_MESH;
    const RWG r = build_rwg(m);
    if (r.N == 0) return BEM_ERR_MESH;
    const BlockPartition P = block_partition(m, r, tri_quadrature(quad_order), radius_mult, max_block_rwg,
                                             overlap_layers);
    int total = 0;
    for (size_t b = 0; b < P.own.size(); b++) {
        if (sizes) {
            sizes[2 * b] = (int)P.own[b].size();
            sizes[2 * b + 1] = (int)P.ext[b].size();
        }
        if (rwg_ext)
            for (int x : P.ext[b]) rwg_ext[total++] = x;
        else
            total += (int)P.ext[b].size();
    }
    *n_ext = total;
    return (int)P.own.size();
}

int bem_precond_apply(bem_ctx* c, int precond, const double* R, int nrhs, double* Zout) {
    if (!c || !R || !Zout || nrhs < 1) return BEM_ERR_ARG;
    KtScope kts{c};
    if (precond == BEM_PRECOND_LU && c->factored < 0)
        return bem_fail(c, BEM_ERR_STATE, "precond_apply: no factors");
    if (precond == BEM_PRECOND_BLOCK && (!c->bj.valid || 2 * c->bj.N != c->n))
        return bem_fail(c, BEM_ERR_STATE, "precond_apply: no block-Jacobi preconditioner for this system");
    if (precond != BEM_PRECOND_LU && precond != BEM_PRECOND_BLOCK)
        return bem_fail(c, BEM_ERR_ARG, "precond_apply: unknown preconditioner");
    (void)hipSetDevice(c->device);
    // R goes through the right-hand-side planes, the result through the solution planes
    if (int rc = load_rhs(c, R, nrhs, "precond_apply")) return rc;
    const size_t work = precond == BEM_PRECOND_LU ? lu_precond_work_bytes(c) : bj_work_bytes(c);
    void* scr = nullptr;
    if (int rc = ctx_scratch(c, work + 256, &scr)) return rc;
    const int rc = precond == BEM_PRECOND_LU ? lu_precond_apply(c, c->Br, c->Bi, c->Xr, c->Xi, scr)
                                             : bj_apply(c, c->Br, c->Bi, c->Xr, c->Xi, scr);
    if (rc) return rc;
    BEM_HIP(c, hipStreamSynchronize(c->stream));
    c->has_X = true;
    c->x_tol = 0.0;
    return store_solution(c, nrhs, Zout);
}

int bem_rhs_planewave(bem_ctx* c, double kr, double ki, double eta, const double* khat, const double* E0, int nrhs,
                      double* B_out) {
    if (!c || !khat || !E0 || nrhs < 1) return bem_fail(c, BEM_ERR_ARG, "rhs: bad arguments");
    if (!c->has_mesh) return bem_fail(c, BEM_ERR_STATE, "rhs: no mesh");
    KtScope kts{c};
    if (c->n != 2 * c->rwg.N) return bem_fail(c, BEM_ERR_STATE, "rhs: system size does not match the mesh");
    (void)hipSetDevice(c->device);
    if (int rc = ctx_ensure_rhs(c, nrhs)) return rc;
    EventTimer tm(c);
    if (int rc = rhs_launch(c, kr, ki, eta, khat, E0, nrhs)) return rc;
    c->t_ms[5] = tm.stop();
    c->has_B = true;
    c->rhs_planewave = true;
    if (B_out) {
        const int n = c->n, ld = c->ldr;
        std::vector<double> re((size_t)n * ld), im((size_t)n * ld);
        BEM_HIP(c, hipMemcpy(re.data(), c->Br, re.size() * sizeof(double), hipMemcpyDeviceToHost));
        BEM_HIP(c, hipMemcpy(im.data(), c->Bi, im.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int j = 0; j < nrhs; j++)
            for (int i = 0; i < n; i++) {
                B_out[2 * ((size_t)j * n + i)] = re[(size_t)i * ld + j];
                B_out[2 * ((size_t)j * n + i) + 1] = im[(size_t)i * ld + j];
            }
    }
    return 0;
}

int bem_farfield(bem_ctx* c, const double* cJ, const double* cM, const double* rh, double kr, double ki, double eta,
                 int n_calls, int n_orient, int ndir, double* Fv) {
    if (!c || !cJ || !cM || !rh || !Fv || n_calls < 1 || n_orient < 1 || ndir < 1)
        return bem_fail(c, BEM_ERR_ARG, "farfield: bad arguments");
    if ((n_calls + 1) / 2 > n_orient) return bem_fail(c, BEM_ERR_ARG, "farfield: n_calls needs ceil(n_calls/2) orientations");
    if (!c->has_mesh) return bem_fail(c, BEM_ERR_STATE, "farfield: no mesh");
    KtScope kts{c};
    (void)hipSetDevice(c->device);
    const int N = c->rwg.N;
    // coefficient planes (n_calls x N each for J and M), directions, output
    const size_t nc = (size_t)n_calls * N;
    const size_t bytes = (4 * nc + (size_t)n_orient * ndir * 3 + (size_t)n_calls * ndir * 6) * sizeof(double);
    void* s = nullptr;
    if (int rc = ctx_scratch(c, bytes, &s)) return rc;
    double* d = static_cast<double*>(s);
    std::vector<double> h(4 * nc);
    for (size_t i = 0; i < nc; i++) {
        h[i] = cJ[2 * i];
        h[nc + i] = cJ[2 * i + 1];
        h[2 * nc + i] = cM[2 * i];
        h[3 * nc + i] = cM[2 * i + 1];
    }
    BEM_HIP(c, hipMemcpy(d, h.data(), 4 * nc * sizeof(double), hipMemcpyHostToDevice));
    double* drh = d + 4 * nc;
    BEM_HIP(c, hipMemcpy(drh, rh, (size_t)n_orient * ndir * 3 * sizeof(double), hipMemcpyHostToDevice));
    double* dfv = drh + (size_t)n_orient * ndir * 3;
    EventTimer tm(c);
    if (int rc = ff_launch(c, d, d + nc, d + 2 * nc, d + 3 * nc, N, drh, kr, ki, eta, n_calls, n_orient, ndir, dfv))
        return rc;
    c->t_ms[4] = tm.stop();
    BEM_HIP(c, hipMemcpy(Fv, dfv, (size_t)n_calls * ndir * 6 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int bem_farfield_noalloc(bem_ctx* c, const double* cJ, const double* cM, const double* rh, double kr, double ki,
                         double eta, int n_calls, int n_orient, int ndir, double* Fv) {
    if (!c || !cJ || !cM || !rh || !Fv || n_calls < 1 || n_orient < 1 || ndir < 1)
        return bem_fail(c, BEM_ERR_ARG, "farfield_noalloc: bad arguments");
    if ((n_calls + 1) / 2 > n_orient)
        return bem_fail(c, BEM_ERR_ARG, "farfield_noalloc: n_calls needs ceil(n_calls/2) orientations");
    if (!c->has_mesh) return bem_fail(c, BEM_ERR_STATE, "farfield_noalloc: no mesh");
    (void)hipSetDevice(c->device);
    return ff_noalloc_launch(c, cJ, cM, rh, kr, ki, eta, n_calls, n_orient, ndir, Fv);
}

int bem_orient_mueller(bem_ctx* c, double kr, double ki, double eta, const double* RT, const double* w, int n_orient,
                       const double* theta, int ntheta, int mirror, double* M_avg, double* M_orient) {
    if (!c || !RT || !w || !theta || !M_avg || n_orient < 1 || ntheta < 1)
        return bem_fail(c, BEM_ERR_ARG, "orient_mueller: bad arguments");
    if (!c->has_X) return bem_fail(c, BEM_ERR_STATE, "orient_mueller: no solution on the device");
    KtScope kts{c};
    if (c->nrhs < 2 * n_orient) return bem_fail(c, BEM_ERR_ARG, "orient_mueller: needs 2 solutions per orientation");
    if (c->n != 2 * c->rwg.N) return bem_fail(c, BEM_ERR_STATE, "orient_mueller: solution does not match the mesh");
    (void)hipSetDevice(c->device);
    EventTimer tm(c);
    if (int rc = mueller_launch(c, kr, ki, eta, RT, w, n_orient, theta, ntheta, mirror, M_avg, M_orient)) return rc;
    c->t_ms[4] = tm.stop();
    c->last_RT.assign(RT, RT + 9 * n_orient);
    c->last_w.assign(w, w + n_orient);
    c->last_theta.assign(theta, theta + ntheta);
    c->last_norient = n_orient;
    c->last_ntheta = ntheta;
    c->last_mirror = mirror;
    c->last_kext[0] = kr;
    c->last_kext[1] = ki;
    c->last_eta_ext = eta;
    return 0;
}

int bem_mueller_grad(bem_ctx* c, const double* dL_dM, double* grad) {
    if (!c || !dL_dM || !grad) return BEM_ERR_ARG;
    if (c->last_norient == 0 || !c->asm_valid || c->factored < 0 || c->factors_stale)
        return bem_fail(c, BEM_ERR_STATE,
                        "mueller_grad: needs assemble, factor (of this matrix), solve and orient_mueller first");
    (void)hipSetDevice(c->device);
    EventTimer tm(c);
    const int rc = adjoint_launch(c, dL_dM, grad);
    c->t_ms[6] = tm.stop();
    return rc;
}

}  // extern "C"
