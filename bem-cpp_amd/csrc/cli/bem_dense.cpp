// bem_dense -- command-line driver for the dense PMCHWT path.
//
// Accepts the dense-path options of the reference CLI (main.cpp:94-181) with
// the same meaning and defaults (main.cpp:60-91) and writes the same JSON
// (output.cpp:5-64).  The pipeline stays on the device: assembly, LU, all
// right-hand sides at once, multi-RHS solve with refinement, far field and
// Mueller average.  Iterative solvers (--solver fmm/pfft/spfft/auto) and their
// options are not part of this library; asking for them is an error.
//
// Extra options:
//   --precision fp32|fp64   factorisation precision (default fp32: FP32 LU on
//                           the matrix cores + FP64 iterative refinement)
//   --ir-tol X --ir-max K   refinement tolerance / sweep limit (1e-10, 10)
//   --device D              GPU index
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/bem_host.h"
#include "bem_abi.h"

namespace {

void usage(const char* p) {
    std::printf("Usage: %s --ka X [options]\n", p);
    std::printf("  --ka X            size parameter (required)\n");
    std::printf("  --ri RE IM        refractive index (default 1.3116 0)\n");
    std::printf("  --ref N           mesh refinements (default 3)\n");
    std::printf("  --shape S         sphere | hex (default sphere)\n");
    std::printf("  --ar X            hex aspect ratio H/D (default 1.0)\n");
    std::printf("  --obj FILE        load an OBJ mesh (normalised to unit equal-volume radius)\n");
    std::printf("  --subdiv N        subdivide the OBJ mesh N times\n");
    std::printf("  --single          one orientation (k along +z)\n");
    std::printf("  --orient A B G    orientation grid (default 8 8 1)\n");
    std::printf("  --orient-sym B G  beta / gamma symmetry factors\n");
    std::printf("  --gamma-mirror    gamma mirror symmetry\n");
    std::printf("  --orient-range I0 I1  orientations [I0, I1) only\n");
    std::printf("  --orient-auto [R] orientation count from ka (default R 4)\n");
    std::printf("  --ntheta N        scattering angles (default 181)\n");
    std::printf("  --quad N          triangle quadrature 4 | 7 | 13 (default 7)\n");
    std::printf("  --out FILE        result file (default result.json)\n");
    std::printf("  --solver dense    the only solver of this build\n");
    std::printf("  --precision P     fp32 (LU on matrix cores + FP64 refinement, default) | fp64\n");
    std::printf("  --ir-tol X        refinement tolerance (default 1e-10)\n");
    std::printf("  --ir-max K        refinement sweeps (default 10)\n");
    std::printf("  --device D        GPU index (default 0)\n");
}

int fail(bem_ctx* c, const char* what, int rc) {
    std::fprintf(stderr, "Error in %s: %s (code %d)\n", what, c ? bem_last_error(c) : "", rc);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    setbuf(stdout, nullptr);
    double ka = 0, n_re = 1.3116, n_im = 0.0;
    int refinements = 3, n_alpha = 8, n_beta = 8, n_gamma = 1, ntheta = 181, quad_order = 7;
    const char* outfile = "result.json";
    bool single = false, gamma_mirror = false, orient_auto = false;
    int beta_sym = 1, gamma_sym = 1, oi0 = -1, oi1 = -1, subdiv = 0, device = 0, ir_max = 10;
    double orient_auto_R = 4.0, hex_ar = 1.0, ir_tol = 1e-10;
    std::string shape = "sphere", precision = "fp32";
    const char* obj = nullptr;
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        auto more = [&](int k) { return i + k < argc; };
        if (!std::strcmp(a, "--ka") && more(1)) ka = std::atof(argv[++i]);
        else if (!std::strcmp(a, "--ri") && more(2)) { n_re = std::atof(argv[++i]); n_im = std::atof(argv[++i]); }
        else if (!std::strcmp(a, "--ref") && more(1)) refinements = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--orient") && more(3)) {
            n_alpha = std::atoi(argv[++i]); n_beta = std::atoi(argv[++i]); n_gamma = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "--ntheta") && more(1)) ntheta = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--quad") && more(1)) quad_order = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--out") && more(1)) outfile = argv[++i];
        else if (!std::strcmp(a, "--single")) single = true;
        else if (!std::strcmp(a, "--solver") && more(1)) {
            const char* s = argv[++i];
            if (std::strcmp(s, "dense") != 0) {
                std::fprintf(stderr, "Solver '%s' is not part of this build (dense only)\n", s);
                return 1;
            }
        } else if (!std::strcmp(a, "--shape") && more(1)) shape = argv[++i];
        else if (!std::strcmp(a, "--ar") && more(1)) hex_ar = std::atof(argv[++i]);
        else if (!std::strcmp(a, "--obj") && more(1)) obj = argv[++i];
        else if (!std::strcmp(a, "--subdiv") && more(1)) subdiv = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--orient-sym") && more(2)) {
            beta_sym = std::atoi(argv[++i]); gamma_sym = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "--gamma-mirror")) gamma_mirror = true;
        else if (!std::strcmp(a, "--orient-range") && more(2)) { oi0 = std::atoi(argv[++i]); oi1 = std::atoi(argv[++i]); }
        else if (!std::strcmp(a, "--orient-auto")) {
            orient_auto = true;
            if (more(1) && argv[i + 1][0] != '-') orient_auto_R = std::atof(argv[++i]);
        } else if (!std::strcmp(a, "--precision") && more(1)) precision = argv[++i];
        else if (!std::strcmp(a, "--ir-tol") && more(1)) ir_tol = std::atof(argv[++i]);
        else if (!std::strcmp(a, "--ir-max") && more(1)) ir_max = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--device") && more(1)) device = std::atoi(argv[++i]);
        else if (!std::strcmp(a, "--help") || !std::strcmp(a, "-h")) { usage(argv[0]); return 0; }
        else {
            std::fprintf(stderr, "Unknown option: %s\n", a);
            usage(argv[0]);
            return 1;
        }
    }
    if (ka <= 0) {
        std::fprintf(stderr, "Error: --ka must be positive\n");
        usage(argv[0]);
        return 1;
    }
    if (precision != "fp32" && precision != "fp64") {
        std::fprintf(stderr, "Error: --precision must be fp32 or fp64\n");
        return 1;
    }
    Timer total;
    const cdouble m(n_re, n_im), k_ext(ka, 0.0), k_int = k_ext * m;  // main.cpp:192-196
    const double eta_ext = 1.0, eta_int = 1.0 / std::abs(m);
    std::printf("=== BEM dense PMCHWT (MI355X) ===\n");
    std::printf("  ka = %.4f, m = %.4f + %.4fi\n", ka, n_re, n_im);
    std::printf("  k_ext = %.4f, k_int = %.4f + %.4fi\n", k_ext.real(), k_int.real(), k_int.imag());
    std::printf("  eta_ext = %.4f, eta_int = %.4f\n", eta_ext, eta_int);
    std::printf("  Refinements: %d, Quad order: %d, precision: %s\n", refinements, quad_order, precision.c_str());

    Mesh mesh;
    if (obj) {
        mesh = load_obj(obj);
        if (mesh.nt() == 0) return 1;
        const double a_eq = normalize_mesh(mesh);
        std::printf("  OBJ normalized: a_eq=%.4f, scaled to unit\n", a_eq);
        for (int s = 0; s < subdiv; s++) mesh = subdivide_flat(mesh);
    } else if (shape == "hex") {
        mesh = hex_prism(hex_ar, refinements);
    } else {
        mesh = icosphere(1.0, refinements);
    }
    const double dmax = mesh_dmax(mesh);
    double area = 0;
    for (int t = 0; t < mesh.nt(); t++) area += mesh.tri_area(t);
    const double h_avg = std::sqrt(2.0 * area / (mesh.nt() * std::sqrt(3.0)));
    const double ppw_int = 2.0 * M_PI / (ka * std::abs(m)) / h_avg;
    std::printf("  Mesh: %d vertices, %d triangles, D_max=%.3f\n", mesh.nv(), mesh.nt(), dmax);
    std::printf("  Resolution: h_avg=%.4f, lambda_ext/h=%.1f, lambda_int/h=%.1f%s\n", h_avg, 2.0 * M_PI / ka / h_avg,
                ppw_int, ppw_int < 6.0 ? " [WARNING: mesh too coarse!]" : ppw_int < 10.0 ? " [marginal]" : " [OK]");
    if (orient_auto && !single) {  // main.cpp:263-277
        const double xi = 0.69 * (2.0 * M_PI / ka) / dmax * (180.0 / M_PI);
        const double delta = xi / 3.0 * orient_auto_R;
        n_alpha = std::max(4, (int)std::ceil(360.0 / delta));
        n_beta = std::max(2, (int)std::ceil(180.0 / delta));
        n_gamma = 1;
        if (n_alpha % 2 == 0) n_alpha++;
        if (n_beta % 2 == 0) n_beta++;
    }

    bem_ctx* c = bem_create(device);
    if (!c) {
        std::fprintf(stderr, "Error: cannot open GPU %d\n", device);
        return 1;
    }
    std::vector<double> verts(3 * (size_t)mesh.nv());
    for (int i = 0; i < mesh.nv(); i++) {
        verts[3 * i] = mesh.verts[i].x;
        verts[3 * i + 1] = mesh.verts[i].y;
        verts[3 * i + 2] = mesh.verts[i].z;
    }
    int rc = bem_set_mesh(c, verts.data(), mesh.nv(), mesh.tris.data(), mesh.nt(), quad_order);
    if (rc) return fail(c, "set_mesh", rc);
    const int N = bem_num_rwg(c);
    std::printf("  RWG: %d basis functions (system %d x %d)\n", N, 2 * N, 2 * N);

    std::vector<double> theta(ntheta);
    for (int i = 0; i < ntheta; i++) theta[i] = M_PI * i / (ntheta - 1);  // main.cpp:331-333

    Timer ta;
    rc = bem_assemble(c, k_ext.real(), k_ext.imag(), k_int.real(), k_int.imag(), eta_ext, eta_int,
                      (unsigned)BEM_ASM_Z64 | (precision == "fp32" ? (unsigned)BEM_ASM_Z32 : 0u));
    if (rc) return fail(c, "assemble", rc);
    const double time_assembly = ta.elapsed_s();

    // orientations and right-hand sides (main.cpp:673-675, 724-751)
    std::vector<Orientation> orients;
    if (single) {
        Orientation o;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) o.RT.m[i][j] = i == j ? 1.0 : 0.0;
        o.weight = 1.0;
        orients.push_back(o);
    } else {
        orients = generate_orientations(n_alpha, n_beta, n_gamma, beta_sym, gamma_sym, gamma_mirror);
        sort_orientations_nearest(orients);
        const int ntot = (int)orients.size();
        int s0 = 0, s1 = ntot;
        if (oi0 >= 0 && oi1 > oi0) {
            s0 = std::min(oi0, ntot);
            s1 = std::min(oi1, ntot);
            std::printf("  Orient range: [%d, %d) of %d\n", s0, s1, ntot);
        }
        orients = std::vector<Orientation>(orients.begin() + s0, orients.begin() + s1);
        std::printf("  Orientations: %d\n", (int)orients.size());
    }
    const int no = (int)orients.size();
    if (no == 0) {
        std::fprintf(stderr, "Error: empty orientation range\n");
        return 1;
    }
    std::vector<double> kh(6 * (size_t)no), e0(6 * (size_t)no), RT(9 * (size_t)no), w(no);
    for (int o = 0; o < no; o++) {
        const Mat3& R = orients[o].RT;
        const Vec3 k = R * Vec3(0, 0, 1), ep = R * Vec3(1, 0, 0), eq = R * Vec3(0, 1, 0);
        const double kk[3] = {k.x, k.y, k.z}, p[3] = {ep.x, ep.y, ep.z}, q[3] = {eq.x, eq.y, eq.z};
        for (int d = 0; d < 3; d++) {
            kh[6 * o + d] = kk[d];
            kh[6 * o + 3 + d] = kk[d];
            e0[6 * o + d] = p[d];
            e0[6 * o + 3 + d] = q[d];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) RT[9 * o + 3 * i + j] = R.m[i][j];
        w[o] = orients[o].weight;
    }
    Timer ts;
    rc = bem_rhs_planewave(c, k_ext.real(), k_ext.imag(), eta_ext, kh.data(), e0.data(), 2 * no, nullptr);
    if (rc) return fail(c, "rhs", rc);
    int prec = precision == "fp32" ? BEM_F32_3M : BEM_F64;
    rc = bem_factor(c, prec);
    if (rc < 0) return fail(c, "factor", rc);
    if (rc > 0) std::fprintf(stderr, "  LU factorization: zero pivot, info=%d\n", rc);
    bem_solve_stats st{};
    rc = bem_solve(c, nullptr, 2 * no, nullptr, ir_tol, prec == BEM_F32_3M ? ir_max : 0, &st);
    if (rc == BEM_ERR_NOCONV) {
        std::printf("  Refinement did not converge (residual %.3e); refactoring in FP64\n", st.residual);
        rc = bem_factor(c, BEM_F64);
        if (rc < 0) return fail(c, "factor", rc);
        rc = bem_solve(c, nullptr, 2 * no, nullptr, 0.0, 0, &st);
    }
    if (rc) return fail(c, "solve", rc);
    if (prec == BEM_F32_3M)
        std::printf("  Refinement: %d sweeps, residual %.3e -> %.3e\n", st.iterations, st.residual0, st.residual);
    const double time_solve = ts.elapsed_s();

    Timer tf;
    std::vector<double> M(16 * (size_t)ntheta), Mo(16 * (size_t)ntheta * no);
    rc = bem_orient_mueller(c, k_ext.real(), k_ext.imag(), eta_ext, RT.data(), w.data(), no, theta.data(), ntheta,
                            gamma_mirror ? 1 : 0, M.data(), Mo.data());
    if (rc) return fail(c, "far field", rc);
    // the per-orientation file holds the matrices before the 1/|k|^2 of the
    // average (main.cpp:554-569, 640-653)
    const double k2 = std::norm(k_ext);
    for (double& v : Mo) v *= k2;
    const double time_farfield = tf.elapsed_s();
    double dev[8];
    bem_timings(c, dev, 8);
    std::printf("  Device time: assembly %.1f ms, factor %.1f ms, solve %.1f ms, refine %.1f ms, far field %.1f ms\n",
                dev[0], dev[1], dev[2], dev[3], dev[4]);
    bem_destroy(c);

    const double time_total = total.elapsed_s();
    const int na_out = single ? 1 : n_alpha, nb_out = single ? 1 : n_beta, ng_out = single ? 1 : n_gamma;
    if (!write_json(outfile, M.data(), theta.data(), ntheta, ka, n_re, n_im, refinements, na_out, nb_out, ng_out,
                    time_assembly, time_solve, time_farfield, time_total))
        return 1;
    const std::string po_path = std::string(outfile) + ".per_orient";
    if (!write_per_orient(po_path.c_str(), no, ntheta, na_out, nb_out, w.data(), Mo.data())) return 1;
    std::printf("\n=== Done ===\n  Assembly: %.1fs\n  Solve: %.1fs\n  Farfield: %.1fs\n  Total: %.1fs\n", time_assembly,
                time_solve, time_farfield, time_total);
    return 0;
}
