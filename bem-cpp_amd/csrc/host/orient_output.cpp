// orient_output.cpp -- orientation quadrature and the result file.
//   euler_rotation             orient.cpp:4-36  (R = Rz(alpha) Ry(beta) Rz(gamma))
//   gauss_legendre             orient.cpp:39-70 (Newton on P_n, nodes descending)
//   generate_orientations      orient.cpp:73-124
//   sort_orientations_nearest  orient.cpp:127-163 (greedy Frobenius nearest neighbour)
//   write_json                 output.cpp:5-64 (byte-compatible format)
//   write_per_orient           main.cpp:640-653 (the .per_orient binary)
#include <cmath>
#include <cstdio>

#include "bem_host.h"

Mat3 euler_rotation(double alpha, double beta, double gamma) {
    const double ca = std::cos(alpha), sa = std::sin(alpha);
    const double cb = std::cos(beta), sb = std::sin(beta);
    const double cg = std::cos(gamma), sg = std::sin(gamma);
    const double A[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    const double B[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    const double C[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
    Mat3 R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) {
                double bc = 0.0;  // (B C)[k][j]
                for (int l = 0; l < 3; l++) bc += B[k][l] * C[l][j];
                s += A[i][k] * bc;
            }
            R.m[i][j] = s;
        }
    return R;
}

// Legendre P_n and P_{n-1} at x by the three-term recurrence.
static void legendre_pair(int n, double x, double& pn, double& pn1) {
    double p_prev = 1.0, p = x;
    for (int j = 2; j <= n; j++) {
        const double next = ((2 * j - 1) * x * p - (j - 1) * p_prev) / j;
        p_prev = p;
        p = next;
    }
    pn = p;
    pn1 = p_prev;
}

void gauss_legendre(int n, std::vector<double>& nodes, std::vector<double>& weights) {
    nodes.assign(n, 0.0);
    weights.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p, q;
            legendre_pair(n, x, p, q);
            const double dp = n * (x * p - q) / (x * x - 1.0);
            const double dx = -p / dp;
            x += dx;
            if (std::fabs(dx) < 1e-15) break;
        }
        double p, q;
        legendre_pair(n, x, p, q);
        const double dp = n * (x * p - q) / (x * x - 1.0);
        nodes[i] = x;
        weights[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

std::vector<Orientation> generate_orientations(int n_alpha, int n_beta, int n_gamma, int beta_sym, int gamma_sym,
                                               bool gamma_mirror) {
    std::vector<double> mu, wmu;
    if (beta_sym == 2) {  // keep the positive half of GL(2 n_beta), doubled weights
        std::vector<double> x, w;
        gauss_legendre(2 * n_beta, x, w);
        for (int i = 0; i < 2 * n_beta; i++)
            if (x[i] > 0) {
                mu.push_back(x[i]);
                wmu.push_back(2.0 * w[i]);
            }
    } else {
        gauss_legendre(n_beta, mu, wmu);
    }
    const double d_alpha = 2.0 * M_PI / n_alpha;
    const double d_gamma = 2.0 * M_PI / n_gamma;  // full-range weight, symmetry folded into the range
    const int gsym = gamma_mirror ? 2 * gamma_sym : gamma_sym;
    const double g_step = 2.0 * M_PI / (gsym * n_gamma);
    std::vector<Orientation> out;
    out.reserve((size_t)n_alpha * n_beta * n_gamma);
    for (int ia = 0; ia < n_alpha; ia++)
        for (int ib = 0; ib < (int)mu.size(); ib++)
            for (int ig = 0; ig < n_gamma; ig++) {
                const double gamma = gamma_mirror ? (ig + 0.5) * g_step : ig * g_step;
                Orientation o;
                o.RT = euler_rotation(ia * d_alpha, std::acos(mu[ib]), gamma).T();
                o.weight = d_alpha * wmu[ib] * d_gamma / (8.0 * M_PI * M_PI);
                out.push_back(o);
            }
    return out;
}

void sort_orientations_nearest(std::vector<Orientation>& orients) {
    const int n = (int)orients.size();
    if (n <= 2) return;
    std::vector<char> taken(n, 0);
    std::vector<Orientation> path;
    path.reserve(n);
    path.push_back(orients[0]);
    taken[0] = 1;
    for (int step = 1; step < n; step++) {
        const Mat3& cur = path.back().RT;
        int best = -1;
        double best_d = 1e30;
        for (int j = 0; j < n; j++) {
            if (taken[j]) continue;
            double d = 0.0;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    const double e = cur.m[r][c] - orients[j].RT.m[r][c];
                    d += e * e;
                }
            if (d < best_d) {
                best_d = d;
                best = j;
            }
        }
        path.push_back(orients[best]);
        taken[best] = 1;
    }
    orients.swap(path);
}

bool write_json(const char* filename, const double* M, const double* theta, int ntheta, double ka, double n_re,
                double n_im, int refinements, int n_alpha, int n_beta, int n_gamma, double time_assembly,
                double time_solve, double time_farfield, double time_total) {
    FILE* f = std::fopen(filename, "w");
    if (!f) {
        std::fprintf(stderr, "Cannot open %s for writing\n", filename);
        return false;
    }
    std::fprintf(f, "{\n  \"ka\": %.6f,\n  \"ri\": [%.6f, %.6f],\n  \"refinements\": %d,\n", ka, n_re, n_im,
                 refinements);
    std::fprintf(f, "  \"orient\": [%d, %d, %d],\n  \"ntheta\": %d,\n", n_alpha, n_beta, n_gamma, ntheta);
    std::fprintf(f, "  \"timing\": {\n    \"assembly_s\": %.2f,\n    \"solve_s\": %.2f,\n", time_assembly,
                 time_solve);
    std::fprintf(f, "    \"farfield_s\": %.2f,\n    \"total_s\": %.2f\n  },\n", time_farfield, time_total);
    std::fprintf(f, "  \"theta\": [");
    for (int t = 0; t < ntheta; t++) std::fprintf(f, t + 1 < ntheta ? "%.4f, " : "%.4f", theta[t] * 180.0 / M_PI);
    std::fprintf(f, "],\n  \"mueller\": [\n");
    for (int i = 0; i < 4; i++) {
        std::fprintf(f, "    [\n");
        for (int j = 0; j < 4; j++) {
            std::fprintf(f, "      [");
            for (int t = 0; t < ntheta; t++)
                std::fprintf(f, t + 1 < ntheta ? "%.8e, " : "%.8e", M[(i * 4 + j) * ntheta + t]);
            std::fprintf(f, j < 3 ? "],\n" : "]\n");
        }
        std::fprintf(f, i < 3 ? "    ],\n" : "    ]\n");
    }
    std::fprintf(f, "  ]\n}\n");
    const bool ok = !std::ferror(f);
    if (std::fclose(f) != 0 || !ok) return false;
    std::printf("Results written to %s\n", filename);
    return true;
}

bool write_per_orient(const char* filename, int n_orient, int ntheta, int n_alpha, int n_beta,
                      const double* weights, const double* M_orient) {
    FILE* f = std::fopen(filename, "wb");
    if (!f) {
        std::fprintf(stderr, "Cannot open %s for writing\n", filename);
        return false;
    }
    const int header[4] = {n_orient, ntheta, n_alpha, n_beta};
    const size_t nm = (size_t)n_orient * 16 * ntheta;
    bool ok = std::fwrite(header, sizeof(int), 4, f) == 4;
    ok = ok && std::fwrite(weights, sizeof(double), (size_t)n_orient, f) == (size_t)n_orient;
    ok = ok && std::fwrite(M_orient, sizeof(double), nm, f) == nm;
    ok = (std::fclose(f) == 0) && ok;
    if (ok) std::printf("  Per-orient Mueller saved: %s (%d orient x 16 x %d)\n", filename, n_orient, ntheta);
    return ok;
}
