// dense_api.cpp -- the reference's dense-path C++ functions (same names,
// arguments, layouts and return values) implemented over the C ABI, so code
// written against pmchwt.h / assembly.h / solver.h / rhs.h / farfield.h links
// against libbem_amd.so unchanged.  All calls share one process-wide device
// context (device from $BEM_DEVICE, default 0), created on first use; like the
// reference they take and return host buffers and block until done.
//
// Errors: the reference exits on CUDA errors (types.h:12-35); these functions
// print the library's message to stderr and exit with status 1 in the same
// situations, and return LAPACK-style info from the factorisations.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bem_abi.h"
#include "bem_host.h"

namespace {

bem_ctx* g_ctx = nullptr;
std::vector<double> g_mesh_key;  // vertices + triangles of the mesh on the device
int g_quad = -1;

bem_ctx* ctx() {
    if (!g_ctx) {
        const char* dev = std::getenv("BEM_DEVICE");
        g_ctx = bem_create(dev ? std::atoi(dev) : 0);
        if (!g_ctx) {
            std::fprintf(stderr, "bem: cannot create a device context\n");
            std::exit(1);
        }
    }
    return g_ctx;
}

void check(int rc, const char* what) {
    if (rc < 0) {
        std::fprintf(stderr, "bem error in %s: %s (code %d)\n", what, bem_last_error(g_ctx), rc);
        std::exit(1);
    }
}

void use_mesh(const Mesh& mesh, int quad_order) {
    std::vector<double> key;
    key.reserve(mesh.verts.size() * 3 + mesh.tris.size() + 1);
    for (const Vec3& v : mesh.verts) key.insert(key.end(), {v.x, v.y, v.z});
    for (int t : mesh.tris) key.push_back(t);
    if (key == g_mesh_key && quad_order == g_quad) return;
    std::vector<double> verts(mesh.verts.size() * 3);
    for (size_t i = 0; i < mesh.verts.size(); i++) {
        verts[3 * i] = mesh.verts[i].x;
        verts[3 * i + 1] = mesh.verts[i].y;
        verts[3 * i + 2] = mesh.verts[i].z;
    }
    check(bem_set_mesh(ctx(), verts.data(), mesh.nv(), mesh.tris.data(), mesh.nt(), quad_order), "set_mesh");
    g_mesh_key.swap(key);
    g_quad = quad_order;
}

}  // namespace

void assemble_L_K_cuda(const RWG& rwg, const Mesh& mesh, cdouble k, int quad_order, cdouble* L, cdouble* K) {
    (void)rwg;  // rebuilt from the mesh in the same (std::map) order
    use_mesh(mesh, quad_order);
    check(bem_assemble_LK(ctx(), k.real(), k.imag(), reinterpret_cast<double*>(L), reinterpret_cast<double*>(K)),
          "assemble_L_K");
}

void assemble_pmchwt(const RWG& rwg, const Mesh& mesh, cdouble k_ext, cdouble k_int, double eta_ext, double eta_int,
                     int quad_order, cdouble* Z, cdouble* L_ext, cdouble* K_ext) {
    if (L_ext && K_ext) assemble_L_K_cuda(rwg, mesh, k_ext, quad_order, L_ext, K_ext);
    use_mesh(mesh, quad_order);
    check(bem_assemble(ctx(), k_ext.real(), k_ext.imag(), k_int.real(), k_int.imag(), eta_ext, eta_int, BEM_ASM_Z64),
          "assemble_pmchwt");
    check(bem_get_Z(ctx(), reinterpret_cast<double*>(Z), 1), "assemble_pmchwt");
    if (L_ext && !K_ext) {
        std::vector<cdouble> k(rwg.N * (size_t)rwg.N);
        assemble_L_K_cuda(rwg, mesh, k_ext, quad_order, L_ext, k.data());
    } else if (K_ext && !L_ext) {
        std::vector<cdouble> l(rwg.N * (size_t)rwg.N);
        assemble_L_K_cuda(rwg, mesh, k_ext, quad_order, l.data(), K_ext);
    }
}

// solver.cu:7-48: the factors and pivots are copied back whatever the info
// (a zero pivot leaves a singular U, as Zgetrf does), and info is returned.
int lu_factorize_cuda(cdouble* Z, int n, int* ipiv) {
    if (n == 0) return 0;
    bem_ctx* c = ctx();
    g_mesh_key.clear();
    check(bem_set_Z(c, reinterpret_cast<const double*>(Z), n, 1), "lu_factorize");
    const int info = bem_factor(c, BEM_F64);
    check(info, "lu_factorize");
    if (info > 0) std::fprintf(stderr, "  LU factorization failed: info=%d\n", info);
    check(bem_get_lu(c, reinterpret_cast<double*>(Z)), "lu_factorize");
    check(bem_get_pivots(c, ipiv), "lu_factorize");
    return info;
}

// solver.cu:51-84: returns the solve's info, which (as Zgetrs's) is -i for an
// illegal i-th argument (n: 2, nrhs: 3) and 0 otherwise.
int lu_solve_cuda(const cdouble* Z, const int* ipiv, int n, cdouble* B, int nrhs) {
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (n == 0 || nrhs == 0) return 0;
    bem_ctx* c = ctx();
    g_mesh_key.clear();
    check(bem_set_lu(c, reinterpret_cast<const double*>(Z), ipiv, n), "lu_solve");
    const int rc =
        bem_solve(c, reinterpret_cast<const double*>(B), nrhs, reinterpret_cast<double*>(B), 0.0, 0, nullptr);
    check(rc, "lu_solve");
    return rc;
}

int lu_solve_full(cdouble* Z, int n, cdouble* B, int nrhs) {
    if (n == 0) return 0;
    bem_ctx* c = ctx();
    g_mesh_key.clear();
    check(bem_set_Z(c, reinterpret_cast<const double*>(Z), n, 1), "lu_solve_full");
    const int info = bem_factor(c, BEM_F64);
    check(info, "lu_solve_full");
    if (info > 0) {  // solver.cu:115-121: B is left untouched
        std::fprintf(stderr, "  LU factorization failed: info=%d\n", info);
        return info;
    }
    check(bem_solve(c, reinterpret_cast<const double*>(B), nrhs, reinterpret_cast<double*>(B), 0.0, 0, nullptr),
          "lu_solve_full");
    check(bem_get_lu(c, reinterpret_cast<double*>(Z)), "lu_solve_full");
    return 0;
}

void compute_rhs_planewave(const RWG& rwg, const Mesh& mesh, cdouble k_ext, double eta_ext, const Vec3& E0,
                           const Vec3& k_hat, int quad_order, cdouble* b) {
    (void)rwg;
    use_mesh(mesh, quad_order);
    const double kh[3] = {k_hat.x, k_hat.y, k_hat.z}, e0[3] = {E0.x, E0.y, E0.z};
    check(bem_rhs_planewave(ctx(), k_ext.real(), k_ext.imag(), eta_ext, kh, e0, 1, reinterpret_cast<double*>(b)),
          "compute_rhs_planewave");
}

void FFCache::init(const RWG& rwg, const Mesh& mesh, int order) {
    const TriQuad q = tri_quadrature(order);
    N = rwg.N;
    Nq = q.npts;
    mesh_ref = &mesh;
    quad_order = order;
    const size_t total = 2 * (size_t)N * Nq;
    qpts.assign(total * 3, 0.0);
    fvals.assign(total * 3, 0.0);
    jw.assign(total, 0.0);
    for (int half = 0; half < 2; half++) {  // farfield.cu:9-52
        const double sign = half == 0 ? 1.0 : -1.0;
        for (int n = 0; n < N; n++) {
            const int t = half == 0 ? rwg.tri_p[n] : rwg.tri_m[n];
            const Vec3 fv = half == 0 ? rwg.free_p[n] : rwg.free_m[n];
            const double A = half == 0 ? rwg.area_p[n] : rwg.area_m[n];
            const double cf = sign * rwg.length[n] / (2.0 * A);
            Vec3 v0, v1, v2;
            mesh.tri_verts(t, v0, v1, v2);
            for (int k = 0; k < Nq; k++) {
                const double l1 = q.pts[k][0], l2 = q.pts[k][1], l0 = 1.0 - l1 - l2;
                const Vec3 r = v0 * l0 + v1 * l1 + v2 * l2;
                const Vec3 f = (r - fv) * cf;
                const size_t i = (size_t)half * N * Nq + (size_t)n * Nq + k;
                qpts[3 * i] = r.x; qpts[3 * i + 1] = r.y; qpts[3 * i + 2] = r.z;
                fvals[3 * i] = f.x; fvals[3 * i + 1] = f.y; fvals[3 * i + 2] = f.z;
                jw[i] = A * q.wts[k];
            }
        }
    }
}

void FFCacheGPU::upload(const FFCache& cache) {
    N = cache.N;
    Nq = cache.Nq;
    host = &cache;
    initialized = true;
    if (cache.mesh_ref) use_mesh(*cache.mesh_ref, cache.quad_order);
}

void FFCacheGPU::free() {
    initialized = false;
    host = nullptr;
}

void compute_farfield_batch_cuda(const FFCacheGPU& gpu_cache, const cdouble* coeffs_J, const cdouble* coeffs_M,
                                 const double* r_hats, cdouble k_ext, double eta_ext, int n_calls, int n_orient,
                                 int ndir, cdouble* Fv_out) {
    if (gpu_cache.host && gpu_cache.host->mesh_ref) use_mesh(*gpu_cache.host->mesh_ref, gpu_cache.host->quad_order);
    check(bem_farfield(ctx(), reinterpret_cast<const double*>(coeffs_J), reinterpret_cast<const double*>(coeffs_M),
                       r_hats, k_ext.real(), k_ext.imag(), eta_ext, n_calls, n_orient, ndir,
                       reinterpret_cast<double*>(Fv_out)),
          "compute_farfield_batch");
}

void compute_far_field(const FFCache& cache, const cdouble* coeffs_J, const cdouble* coeffs_M, cdouble k_ext,
                       double eta_ext, const double* theta_arr, int ntheta, cdouble* F_theta, cdouble* F_phi) {
    if (cache.mesh_ref) use_mesh(*cache.mesh_ref, cache.quad_order);
    std::vector<double> rh(3 * (size_t)ntheta);
    for (int t = 0; t < ntheta; t++) {  // phi = 0 plane, farfield.cu:443-459
        rh[3 * t] = std::sin(theta_arr[t]);
        rh[3 * t + 1] = 0.0;
        rh[3 * t + 2] = std::cos(theta_arr[t]);
    }
    std::vector<cdouble> Fv(3 * (size_t)ntheta);
    check(bem_farfield(ctx(), reinterpret_cast<const double*>(coeffs_J), reinterpret_cast<const double*>(coeffs_M),
                       rh.data(), k_ext.real(), k_ext.imag(), eta_ext, 1, 1, ntheta, reinterpret_cast<double*>(Fv.data())),
          "compute_far_field");
    for (int t = 0; t < ntheta; t++) {
        const double ct = std::cos(theta_arr[t]), st = std::sin(theta_arr[t]);
        F_theta[t] = Fv[3 * t] * ct + Fv[3 * t + 2] * (-st);
        F_phi[t] = Fv[3 * t + 1];
    }
}

void amplitude_to_mueller(const cdouble* S1, const cdouble* S2, const cdouble* S3, const cdouble* S4, int ntheta,
                          double* M) {
    // Bohren & Huffman eq. 3.16 in the element order of farfield.cu:464-508
    for (int t = 0; t < ntheta; t++) {
        const double a1 = std::norm(S1[t]), a2 = std::norm(S2[t]), a3 = std::norm(S3[t]), a4 = std::norm(S4[t]);
        const cdouble s23 = S2[t] * std::conj(S3[t]), s14 = S1[t] * std::conj(S4[t]);
        const cdouble s24 = S2[t] * std::conj(S4[t]), s13 = S1[t] * std::conj(S3[t]);
        const cdouble s12 = S1[t] * std::conj(S2[t]), s34 = S3[t] * std::conj(S4[t]);
        const double v[16] = {0.5 * (a1 + a2 + a3 + a4), 0.5 * (a2 - a1 + a4 - a3), s23.real() + s14.real(),
                              s23.imag() - s14.imag(),   0.5 * (a2 - a1 - a4 + a3), 0.5 * (a2 + a1 - a4 - a3),
                              s23.real() - s14.real(),   s23.imag() + s14.imag(),   s24.real() + s13.real(),
                              s24.real() - s13.real(),   s12.real() + s34.real(),   -s12.imag() - s34.imag(),
                              -s24.imag() + s13.imag(),  -s24.imag() - s13.imag(),  s12.imag() - s34.imag(),
                              s12.real() - s34.real()};
        for (int e = 0; e < 16; e++) M[e * ntheta + t] = v[e];
    }
}
