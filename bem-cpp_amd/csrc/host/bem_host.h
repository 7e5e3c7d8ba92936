// bem_host.h -- host-side types and functions of the dense path, with the
// names and argument meaning of the reference's C++ interface so that its
// dense branch (main.cpp:660-854) compiles against this library unchanged.
//
//   Vec3, Mat3, cdouble                types.h:38-66, orient.h:8-24
//   Mesh, icosphere, hex_prism, ...    mesh.h:7-44
//   RWG, build_rwg                     rwg.h:7-18
//   TriQuad, tri_quadrature            quadrature.h:8-70
//   Orientation, euler_rotation, ...   orient.h:26-47
//   write_json                         output.h:9-14
//   assemble_pmchwt, assemble_L_K_cuda pmchwt.h:12-19, assembly.h:11-13
//   lu_factorize_cuda, lu_solve_cuda,
//   lu_solve_full                      solver.h:11-22
//   compute_rhs_planewave              rhs.h:11-16
//   FFCache, FFCacheGPU,
//   compute_farfield_batch_cuda,
//   compute_far_field,
//   amplitude_to_mueller               farfield.h:9-71
//
// No HIP or torch types appear here; the device work is reached through the
// C ABI of include/bem_abi.h.
#pragma once

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

typedef std::complex<double> cdouble;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
static const double INV4PI = 1.0 / (4.0 * M_PI);

struct Vec3 {
    double x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3 cross(const Vec3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    double norm2() const { return x * x + y * y + z * z; }
    Vec3 normalized() const { double n = norm(); return {x / n, y / n, z / n}; }
};
inline Vec3 operator*(double s, const Vec3& v) { return v * s; }

struct Timer {
    std::chrono::steady_clock::time_point t0;
    Timer() { reset(); }
    void reset() { t0 = std::chrono::steady_clock::now(); }
    double elapsed_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    double elapsed_s() const { return elapsed_ms() * 1e-3; }
};

// ---- mesh ------------------------------------------------------------------
struct Mesh {
    std::vector<Vec3> verts;
    std::vector<int> tris;  // flat [v0, v1, v2] per triangle
    int nv() const { return (int)verts.size(); }
    int nt() const { return (int)tris.size() / 3; }
    void tri_verts(int ti, Vec3& a, Vec3& b, Vec3& c) const {
        a = verts[tris[3 * ti]];
        b = verts[tris[3 * ti + 1]];
        c = verts[tris[3 * ti + 2]];
    }
    double tri_area(int ti) const {
        Vec3 a, b, c;
        tri_verts(ti, a, b, c);
        return 0.5 * (b - a).cross(c - a).norm();
    }
};

Mesh icosphere(double radius, int refinements);
Mesh hex_prism(double aspect_ratio, int refinements);
Mesh subdivide_flat(const Mesh& m);
Mesh load_obj(const char* filename);
double mesh_volume(const Mesh& m);
double normalize_mesh(Mesh& m);
double mesh_dmax(const Mesh& m);

// ---- RWG basis -------------------------------------------------------------
struct RWG {
    int N = 0;
    std::vector<int> tri_p, tri_m;
    std::vector<Vec3> free_p, free_m;
    std::vector<double> length, area_p, area_m;
};
RWG build_rwg(const Mesh& mesh);

// ---- triangle quadrature ---------------------------------------------------
struct TriQuad {
    int npts;
    double pts[25][2];  // barycentric (lambda1, lambda2)
    double wts[25];     // sum to one
};
TriQuad tri_quadrature(int order);

// ---- block partition (block-Jacobi / RAS preconditioner) -------------------
struct BlockPartition {
    std::vector<std::vector<int>> own;  // RWG indices of each block, sorted
    std::vector<std::vector<int>> ext;  // own first, then the overlap RWGs, sorted
    double cell_size = 0.0, avg_extent = 0.0;
};
BlockPartition block_partition(const Mesh& mesh, const RWG& rwg, const TriQuad& quad, double radius_mult,
                               int max_block_rwg, int overlap_layers);

// ---- orientations ----------------------------------------------------------
struct Mat3 {
    double m[3][3];
    Vec3 operator*(const Vec3& v) const {
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
    Mat3 T() const {
        Mat3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.m[i][j] = m[j][i];
        return r;
    }
};
struct Orientation {
    Mat3 RT;  // inverse rotation, lab -> body
    double weight;
};
Mat3 euler_rotation(double alpha, double beta, double gamma);
void gauss_legendre(int n, std::vector<double>& nodes, std::vector<double>& weights);
std::vector<Orientation> generate_orientations(int n_alpha, int n_beta, int n_gamma, int beta_sym = 1,
                                               int gamma_sym = 1, bool gamma_mirror = false);
void sort_orientations_nearest(std::vector<Orientation>& orients);

// ---- output ----------------------------------------------------------------
bool write_json(const char* filename, const double* M, const double* theta, int ntheta, double ka, double n_re,
                double n_im, int refinements, int n_alpha, int n_beta, int n_gamma, double time_assembly,
                double time_solve, double time_farfield, double time_total);
// The per-orientation Mueller file of main.cpp:640-653; returns false when
// the file cannot be written.
bool write_per_orient(const char* filename, int n_orient, int ntheta, int n_alpha, int n_beta,
                      const double* weights, const double* M_orient);

// ---- dense path over the device library (dense_api.cpp) --------------------
void assemble_pmchwt(const RWG& rwg, const Mesh& mesh, cdouble k_ext, cdouble k_int, double eta_ext,
                     double eta_int, int quad_order, cdouble* Z, cdouble* L_ext, cdouble* K_ext);
void assemble_L_K_cuda(const RWG& rwg, const Mesh& mesh, cdouble k, int quad_order, cdouble* L, cdouble* K);
int lu_factorize_cuda(cdouble* Z, int n, int* ipiv);
int lu_solve_cuda(const cdouble* Z, const int* ipiv, int n, cdouble* B, int nrhs);
int lu_solve_full(cdouble* Z, int n, cdouble* B, int nrhs);
void compute_rhs_planewave(const RWG& rwg, const Mesh& mesh, cdouble k_ext, double eta_ext, const Vec3& E0,
                           const Vec3& k_hat, int quad_order, cdouble* b);

struct FFCache {
    std::vector<double> qpts, fvals, jw;  // (2 N Nq, 3), (2 N Nq, 3), (2 N Nq)
    int N = 0, Nq = 0;
    void init(const RWG& rwg, const Mesh& mesh, int quad_order);
    const Mesh* mesh_ref = nullptr;  // the mesh the device far field is evaluated on
    int quad_order = 7;
};
struct FFCacheGPU {
    int N = 0, Nq = 0;
    bool initialized = false;
    const FFCache* host = nullptr;
    void upload(const FFCache& cache);
    void free();
    ~FFCacheGPU() { free(); }
};
void compute_farfield_batch_cuda(const FFCacheGPU& gpu_cache, const cdouble* coeffs_J, const cdouble* coeffs_M,
                                 const double* r_hats, cdouble k_ext, double eta_ext, int n_calls, int n_orient,
                                 int ndir, cdouble* Fv_out);
void compute_far_field(const FFCache& cache, const cdouble* coeffs_J, const cdouble* coeffs_M, cdouble k_ext,
                       double eta_ext, const double* theta_arr, int ntheta, cdouble* F_theta, cdouble* F_phi);
void amplitude_to_mueller(const cdouble* S1, const cdouble* S2, const cdouble* S3, const cdouble* S4, int ntheta,
                          double* M);
