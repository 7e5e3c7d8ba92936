// ref_api_dense -- a caller written against the reference's own C++ dense-path
// interface (bem_host.h: pmchwt.h, solver.h, rhs.h, farfield.h, orient.h),
// making exactly the calls of the reference's dense branch
// (main.cpp:660-846) in the same order, so that the drop-in claim is
// exercised by a compiled program and not only through the C ABI:
//
//   single orientation (main.cpp:670-719): assemble_pmchwt, lu_factorize_cuda,
//     compute_rhs_planewave x 2, lu_solve_cuda, compute_far_field x 2,
//     amplitude_to_mueller;
//   orientation average (main.cpp:720-846): generate_orientations,
//     sort_orientations_nearest, compute_rhs_planewave per orientation and
//     polarisation, lu_solve_full, FFCacheGPU::upload,
//     compute_farfield_batch_cuda, amplitude_to_mueller, weighted sum.
//
// Usage:
//   ref_api_dense <sphere|hex> <refinements> <param> <ka> <m_re> <m_im>
//                 <n_alpha> <n_beta> <n_gamma> <ntheta> <out.bin>
//     (n_alpha = 0: the single-orientation branch); writes the Mueller
//     matrix M[(i*4+j)*ntheta + t] as 16 * ntheta raw doubles.
//   ref_api_dense shims
//     checks the return conventions of the solver shims (solver.cu:7-84):
//     a zero column gives info > 0 with factors and pivots still copied back,
//     and lu_solve_cuda returns the solve's info; exit status 0 when they hold.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/bem_host.h"

namespace {

int dense_branch(int argc, char** argv) {
    if (argc != 12) {
        std::fprintf(stderr, "usage: %s <sphere|hex> <ref> <param> <ka> <m_re> <m_im> <na> <nb> <ng> <ntheta> "
                             "<out.bin>\n", argv[0]);
        return 2;
    }
    const std::string shape = argv[1];
    const int refinements = std::atoi(argv[2]);
    const double param = std::atof(argv[3]), ka = std::atof(argv[4]);
    const double n_re = std::atof(argv[5]), n_im = std::atof(argv[6]);
    const int n_alpha = std::atoi(argv[7]), n_beta = std::atoi(argv[8]), n_gamma = std::atoi(argv[9]);
    const int ntheta = std::atoi(argv[10]);
    const char* outfile = argv[11];
    const int quad_order = 7;

    // main.cpp:116-196: mesh, basis, wavenumbers, impedances, angles
    Mesh mesh = shape == "hex" ? hex_prism(param, refinements) : icosphere(param, refinements);
    RWG rwg = build_rwg(mesh);
    const int N = rwg.N, N2 = 2 * N;
    const cdouble m_rel(n_re, n_im);
    const cdouble k_ext(ka, 0.0), k_int = k_ext * m_rel;
    const double eta_ext = 1.0, eta_int = 1.0 / std::abs(m_rel);
    std::vector<double> theta_arr(ntheta);
    for (int i = 0; i < ntheta; i++) theta_arr[i] = M_PI * i / (ntheta - 1);
    FFCache ff_cache;
    ff_cache.init(rwg, mesh, quad_order);
    std::vector<double> M_avg(16 * ntheta, 0.0);

    std::vector<cdouble> Z((size_t)N2 * N2);
    assemble_pmchwt(rwg, mesh, k_ext, k_int, eta_ext, eta_int, quad_order, Z.data(), NULL, NULL);

    if (n_alpha == 0) {
        Vec3 k_hat(0, 0, 1), E_par(1, 0, 0), E_perp(0, 1, 0);
        std::vector<int> ipiv(N2);
        const int info = lu_factorize_cuda(Z.data(), N2, ipiv.data());
        if (info != 0) return 3;
        std::vector<cdouble> B((size_t)N2 * 2);
        compute_rhs_planewave(rwg, mesh, k_ext, eta_ext, E_par, k_hat, quad_order, &B[0]);
        compute_rhs_planewave(rwg, mesh, k_ext, eta_ext, E_perp, k_hat, quad_order, &B[N2]);
        if (lu_solve_cuda(Z.data(), ipiv.data(), N2, B.data(), 2) != 0) return 3;
        std::vector<cdouble> Fth_par(ntheta), Fph_par(ntheta), Fth_perp(ntheta), Fph_perp(ntheta);
        compute_far_field(ff_cache, &B[0], &B[N], k_ext, eta_ext, theta_arr.data(), ntheta, Fth_par.data(),
                          Fph_par.data());
        compute_far_field(ff_cache, &B[N2], &B[N2 + N], k_ext, eta_ext, theta_arr.data(), ntheta, Fth_perp.data(),
                          Fph_perp.data());
        const cdouble ik = cdouble(0, -1) * k_ext;
        std::vector<cdouble> S1(ntheta), S2(ntheta), S3(ntheta), S4(ntheta);
        for (int t = 0; t < ntheta; t++) {
            S2[t] = ik * Fth_par[t];
            S4[t] = ik * Fph_par[t];
            S3[t] = ik * Fth_perp[t];
            S1[t] = ik * Fph_perp[t];
        }
        amplitude_to_mueller(S1.data(), S2.data(), S3.data(), S4.data(), ntheta, M_avg.data());
        const double k2 = std::norm(k_ext);
        for (double& v : M_avg) v /= k2;
    } else {
        std::vector<Orientation> orients = generate_orientations(n_alpha, n_beta, n_gamma);
        sort_orientations_nearest(orients);
        const int n_range = (int)orients.size(), n_calls = 2 * n_range;
        std::vector<cdouble> B((size_t)N2 * n_calls, 0);
        for (int ri = 0; ri < n_range; ri++) {
            const Mat3& RT = orients[ri].RT;
            const Vec3 k_hat = RT * Vec3(0, 0, 1), e_par = RT * Vec3(1, 0, 0), e_perp = RT * Vec3(0, 1, 0);
            compute_rhs_planewave(rwg, mesh, k_ext, eta_ext, e_par, k_hat, quad_order, &B[(size_t)ri * 2 * N2]);
            compute_rhs_planewave(rwg, mesh, k_ext, eta_ext, e_perp, k_hat, quad_order, &B[(size_t)(ri * 2 + 1) * N2]);
        }
        if (lu_solve_full(Z.data(), N2, B.data(), n_calls) != 0) return 3;
        FFCacheGPU ff_gpu;
        ff_gpu.upload(ff_cache);
        std::vector<cdouble> all_J((size_t)n_calls * N), all_M((size_t)n_calls * N);
        for (int c = 0; c < n_calls; c++) {
            std::memcpy(&all_J[(size_t)c * N], &B[(size_t)c * N2], N * sizeof(cdouble));
            std::memcpy(&all_M[(size_t)c * N], &B[(size_t)c * N2 + N], N * sizeof(cdouble));
        }
        std::vector<double> r_hats((size_t)n_range * ntheta * 3);
        std::vector<Vec3> e_th((size_t)n_range * ntheta), e_ph((size_t)n_range * ntheta);
        for (int ri = 0; ri < n_range; ri++)
            for (int it = 0; it < ntheta; it++) {
                const double ct = std::cos(theta_arr[it]), st = std::sin(theta_arr[it]);
                const Mat3& RT = orients[ri].RT;
                const Vec3 rh = RT * Vec3(st, 0, ct);
                const size_t b = ((size_t)ri * ntheta + it) * 3;
                r_hats[b] = rh.x;
                r_hats[b + 1] = rh.y;
                r_hats[b + 2] = rh.z;
                e_th[(size_t)ri * ntheta + it] = RT * Vec3(ct, 0, -st);
                e_ph[(size_t)ri * ntheta + it] = RT * Vec3(0, 1, 0);
            }
        std::vector<cdouble> Fv((size_t)n_calls * ntheta * 3);
        compute_farfield_batch_cuda(ff_gpu, all_J.data(), all_M.data(), r_hats.data(), k_ext, eta_ext, n_calls, n_range,
                                    ntheta, Fv.data());
        const cdouble ik = cdouble(0, -1) * k_ext;
        const double k2 = std::norm(k_ext);
        std::vector<double> M_o(16 * ntheta);
        for (int ri = 0; ri < n_range; ri++) {
            const cdouble* Fp = &Fv[(size_t)(2 * ri) * ntheta * 3];
            const cdouble* Fq = &Fv[(size_t)(2 * ri + 1) * ntheta * 3];
            std::vector<cdouble> S1(ntheta), S2(ntheta), S3(ntheta), S4(ntheta);
            for (int it = 0; it < ntheta; it++) {
                const Vec3& a = e_th[(size_t)ri * ntheta + it];
                const Vec3& b = e_ph[(size_t)ri * ntheta + it];
                S2[it] = ik * (Fp[it * 3] * a.x + Fp[it * 3 + 1] * a.y + Fp[it * 3 + 2] * a.z);
                S4[it] = ik * (Fp[it * 3] * b.x + Fp[it * 3 + 1] * b.y + Fp[it * 3 + 2] * b.z);
                S3[it] = ik * (Fq[it * 3] * a.x + Fq[it * 3 + 1] * a.y + Fq[it * 3 + 2] * a.z);
                S1[it] = ik * (Fq[it * 3] * b.x + Fq[it * 3 + 1] * b.y + Fq[it * 3 + 2] * b.z);
            }
            amplitude_to_mueller(S1.data(), S2.data(), S3.data(), S4.data(), ntheta, M_o.data());
            for (int i = 0; i < 16 * ntheta; i++) M_avg[i] += orients[ri].weight * M_o[i] / k2;
        }
    }
    FILE* f = std::fopen(outfile, "wb");
    if (!f || std::fwrite(M_avg.data(), sizeof(double), M_avg.size(), f) != M_avg.size()) {
        std::fprintf(stderr, "cannot write %s\n", outfile);
        if (f) std::fclose(f);
        return 4;
    }
    std::fclose(f);
    return 0;
}

int shim_checks() {
    int fails = 0;
    auto expect = [&](bool ok, const char* what) {
        std::printf("%s: %s\n", ok ? "ok" : "FAILED", what);
        if (!ok) fails++;
    };
    // a 40 x 40 system whose column 7 is zero: Zgetrf reports info = 8 and
    // still leaves its factors and pivots
    const int n = 40, zc = 7;
    std::vector<cdouble> Z((size_t)n * n), Z0;
    unsigned s = 12345u;
    auto rnd = [&]() {
        s = s * 1664525u + 1013904223u;
        return (double)(s >> 8) / 16777216.0 - 0.5;
    };
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Z[(size_t)i * n + j] = j == zc ? cdouble(0, 0) : cdouble(rnd(), rnd());
    Z0 = Z;
    std::vector<int> ipiv(n, -7);
    const int info = lu_factorize_cuda(Z.data(), n, ipiv.data());
    expect(info == zc + 1, "lu_factorize_cuda returns info = first zero column + 1");
    bool piv_ok = true;
    for (int i = 0; i < n; i++) piv_ok = piv_ok && ipiv[i] >= i + 1 && ipiv[i] <= n;
    expect(piv_ok, "pivots copied back after a zero pivot");
    expect(std::memcmp(Z.data(), Z0.data(), Z.size() * sizeof(cdouble)) != 0, "factors copied back after a zero pivot");
    // the leading zc columns of the factors are exact: L U reproduces the
    // leading zc x zc block of P Z
    std::vector<cdouble> PZ = Z0;
    for (int i = 0; i < n; i++)
        if (ipiv[i] - 1 != i)
            for (int j = 0; j < n; j++) std::swap(PZ[(size_t)i * n + j], PZ[(size_t)(ipiv[i] - 1) * n + j]);
    double err = 0.0, scale = 0.0;
    for (int i = 0; i < zc; i++)
        for (int j = 0; j < zc; j++) {
            cdouble a = 0.0;
            for (int k = 0; k <= std::min(i, j); k++)
                a += (k == i ? cdouble(1, 0) : Z[(size_t)i * n + k]) * Z[(size_t)k * n + j];
            err = std::max(err, std::abs(a - PZ[(size_t)i * n + j]));
            scale = std::max(scale, std::abs(PZ[(size_t)i * n + j]));
        }
    expect(err <= 1e-13 * scale, "factors before the zero pivot reproduce P Z");
    // lu_solve_cuda returns Zgetrs's info: 0, or -i for the illegal i-th argument
    std::vector<cdouble> A((size_t)n * n), B(n), X;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = cdouble(rnd(), rnd()) + (i == j ? cdouble(n, 0) : 0.0);
        B[i] = cdouble(rnd(), rnd());
    }
    std::vector<cdouble> LU = A;
    std::vector<int> ip(n);
    expect(lu_factorize_cuda(LU.data(), n, ip.data()) == 0, "lu_factorize_cuda returns 0 on a regular matrix");
    X = B;
    expect(lu_solve_cuda(LU.data(), ip.data(), n, X.data(), 1) == 0, "lu_solve_cuda returns 0");
    double res = 0.0, bn = 0.0;
    for (int i = 0; i < n; i++) {
        cdouble r = B[i];
        for (int j = 0; j < n; j++) r -= A[(size_t)i * n + j] * X[j];
        res = std::max(res, std::abs(r));
        bn = std::max(bn, std::abs(B[i]));
    }
    expect(res <= 1e-12 * bn, "lu_solve_cuda solves Z X = B");
    expect(lu_solve_cuda(LU.data(), ip.data(), -1, X.data(), 1) == -2, "lu_solve_cuda: n < 0 gives info -2");
    expect(lu_solve_cuda(LU.data(), ip.data(), n, X.data(), -1) == -3, "lu_solve_cuda: nrhs < 0 gives info -3");
    return fails == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "shims") == 0) return shim_checks();
    return dense_branch(argc, argv);
}
