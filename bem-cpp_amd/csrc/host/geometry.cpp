// geometry.cpp -- meshes, RWG basis and triangle quadrature on the host.
//
// The vertex, triangle and edge orders must equal the reference's exactly:
// they fix the row/column order of Z and the solution vectors.
//   icosphere       mesh.cpp:11-67   (normalised midpoint refinement)
//   subdivide_flat  mesh.cpp:70-100  (1 -> 4 split, no projection)
//   hex_prism       mesh.cpp:102-154 (unit equal-volume radius)
//   build_rwg       rwg.cpp:5-43     (edges in (min,max) vertex-pair order)
//   tri_quadrature  quadrature.h:14-70 (Dunavant constants as published there)
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>

#include "bem_host.h"

namespace {

// Midpoint refinement shared by both generators.  Every edge gets one new
// vertex, created the first time the edge is met in triangle order; each
// triangle (a, b, c) becomes (a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca).
template <class MidFn>
void split_all(std::vector<Vec3>& verts, std::vector<int>& tris, MidFn make_mid) {
    std::map<std::pair<int, int>, int> mid_of;
    auto mid = [&](int a, int b) {
        std::pair<int, int> key(std::min(a, b), std::max(a, b));
        auto it = mid_of.find(key);
        if (it != mid_of.end()) return it->second;
        int id = (int)verts.size();
        verts.push_back(make_mid(verts[a], verts[b]));
        mid_of.emplace(key, id);
        return id;
    };
    std::vector<int> out;
    out.reserve(tris.size() * 4);
    for (size_t t = 0; t + 2 < tris.size(); t += 3) {
        int a = tris[t], b = tris[t + 1], c = tris[t + 2];
        int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
        const int quad[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
        out.insert(out.end(), quad, quad + 12);
    }
    tris.swap(out);
}

}  // namespace

Mesh icosphere(double radius, int refinements) {
    const double g = 0.5 * (1.0 + std::sqrt(5.0));
    Mesh m;
    m.verts = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
               {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    const double s = 1.0 / m.verts[0].norm();
    for (Vec3& v : m.verts) v = v * s;
    m.tris = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
              3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int r = 0; r < refinements; r++)
        split_all(m.verts, m.tris, [](const Vec3& a, const Vec3& b) { return ((a + b) * 0.5).normalized(); });
    for (Vec3& v : m.verts) v = v * radius;
    return m;
}

Mesh subdivide_flat(const Mesh& in) {
    Mesh m = in;
    split_all(m.verts, m.tris, [](const Vec3& a, const Vec3& b) { return (a + b) * 0.5; });
    return m;
}

Mesh hex_prism(double aspect_ratio, int refinements) {
    // circumradius a from (3 sqrt3 / 2) a^2 H = 4 pi / 3 with H = AR * a * sqrt3
    const double a = std::pow(8.0 * M_PI / (27.0 * aspect_ratio), 1.0 / 3.0);
    const double H = aspect_ratio * a * std::sqrt(3.0);
    Mesh m;
    m.verts.resize(14);
    m.verts[0] = Vec3(0, 0, 0.5 * H);
    m.verts[7] = Vec3(0, 0, -0.5 * H);
    for (int i = 0; i < 6; i++) {
        const double ang = 2.0 * M_PI * i / 6.0;
        m.verts[1 + i] = Vec3(a * std::cos(ang), a * std::sin(ang), 0.5 * H);
        m.verts[8 + i] = Vec3(a * std::cos(ang), a * std::sin(ang), -0.5 * H);
    }
    for (int i = 0; i < 6; i++) {
        const int t1 = 1 + i, t2 = 1 + (i + 1) % 6, b1 = 8 + i, b2 = 8 + (i + 1) % 6;
        const int f[12] = {0, t1, t2, 7, b2, b1, t1, b1, b2, t1, b2, t2};  // top, bottom, two side triangles
        m.tris.insert(m.tris.end(), f, f + 12);
    }
    for (int r = 0; r < refinements; r++) m = subdivide_flat(m);
    return m;
}

Mesh load_obj(const char* filename) {
    Mesh m;
    std::ifstream f(filename);
    if (!f) {
        std::fprintf(stderr, "Error: cannot open OBJ file: %s\n", filename);
        return m;
    }
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            double x, y, z;
            ss >> x >> y >> z;
            m.verts.emplace_back(x, y, z);
        } else if (tag == "f") {
            std::vector<int> poly;
            std::string tok;
            while (ss >> tok) poly.push_back(std::atoi(tok.c_str()) - 1);
            for (size_t i = 1; i + 1 < poly.size(); i++) {
                m.tris.push_back(poly[0]);
                m.tris.push_back(poly[i]);
                m.tris.push_back(poly[i + 1]);
            }
        }
    }
    std::vector<int> keep;
    for (int t = 0; t < m.nt(); t++) {
        int a = m.tris[3 * t], b = m.tris[3 * t + 1], c = m.tris[3 * t + 2];
        if (a == b || b == c || a == c || m.tri_area(t) < 1e-10) continue;
        keep.insert(keep.end(), {a, b, c});
    }
    m.tris.swap(keep);
    return m;
}

double mesh_volume(const Mesh& m) {
    double v = 0.0;
    for (int t = 0; t < m.nt(); t++) {
        Vec3 a, b, c;
        m.tri_verts(t, a, b, c);
        v += a.dot(b.cross(c));
    }
    return v / 6.0;
}

double normalize_mesh(Mesh& m) {
    double a_eq = std::cbrt(3.0 * std::fabs(mesh_volume(m)) / (4.0 * M_PI));
    if (a_eq < 1e-30) {
        Vec3 c;
        for (const Vec3& v : m.verts) c = c + v;
        c = c * (1.0 / m.nv());
        double r2 = 0;
        for (const Vec3& v : m.verts) r2 = std::max(r2, (v - c).norm2());
        a_eq = std::sqrt(r2);
    }
    for (Vec3& v : m.verts) v = v * (1.0 / a_eq);
    return a_eq;
}

double mesh_dmax(const Mesh& m) {
    Vec3 lo(1e30, 1e30, 1e30), hi(-1e30, -1e30, -1e30);
    for (const Vec3& v : m.verts) {
        lo = Vec3(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
        hi = Vec3(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
    }
    return (hi - lo).norm();
}

RWG build_rwg(const Mesh& mesh) {
    // edge (min, max) -> triangles containing it, with the vertex opposite the edge
    std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> edges;
    for (int t = 0; t < mesh.nt(); t++)
        for (int j = 0; j < 3; j++) {
            const int a = mesh.tris[3 * t + j], b = mesh.tris[3 * t + (j + 1) % 3];
            const int opp = mesh.tris[3 * t + (j + 2) % 3];
            edges[{std::min(a, b), std::max(a, b)}].emplace_back(t, opp);
        }
    RWG r;
    for (const auto& e : edges) {
        if (e.second.size() != 2) continue;  // boundary or non-manifold edge: no basis function
        const int tp = e.second[0].first, tm = e.second[1].first;
        r.tri_p.push_back(tp);
        r.tri_m.push_back(tm);
        r.free_p.push_back(mesh.verts[e.second[0].second]);
        r.free_m.push_back(mesh.verts[e.second[1].second]);
        r.length.push_back((mesh.verts[e.first.second] - mesh.verts[e.first.first]).norm());
        r.area_p.push_back(mesh.tri_area(tp));
        r.area_m.push_back(mesh.tri_area(tm));
    }
    r.N = (int)r.tri_p.size();
    return r;
}

TriQuad tri_quadrature(int order) {
    TriQuad q;
    std::memset(&q, 0, sizeof(q));
    auto put = [&q](int i, double l1, double l2, double w) {
        q.pts[i][0] = l1;
        q.pts[i][1] = l2;
        q.wts[i] = w;
    };
    if (order == 4) {
        q.npts = 4;
        put(0, 1.0 / 3, 1.0 / 3, -27.0 / 48);
        put(1, 0.6, 0.2, 25.0 / 48);
        put(2, 0.2, 0.6, 25.0 / 48);
        put(3, 0.2, 0.2, 25.0 / 48);
    } else if (order == 13) {
        q.npts = 13;
        const double r1 = 0.260345966079038, s1 = 0.479308067841923, w1 = 0.175615257433204;
        const double r2 = 0.065130102902216, s2 = 0.869739794195568, w2 = 0.053347235608839;
        const double r3 = 0.048690315425316, s3 = 0.312865496004875, u3 = 0.638444188569809;
        const double w3 = 0.077113760890257;
        put(0, 1.0 / 3, 1.0 / 3, -0.149570044467670);
        put(1, r1, r1, w1); put(2, r1, s1, w1); put(3, s1, r1, w1);
        put(4, r2, r2, w2); put(5, r2, s2, w2); put(6, s2, r2, w2);
        put(7, r3, s3, w3); put(8, s3, r3, w3); put(9, u3, s3, w3);
        put(10, s3, u3, w3); put(11, r3, u3, w3); put(12, u3, r3, w3);
    } else {  // 7 points, also the fallback for unknown orders
        q.npts = 7;
        const double a1 = 0.059715871789770, a2 = 0.470142064105115;
        const double b1 = 0.797426985353087, b2 = 0.101286507323456;
        const double wa = 0.13239415278851, wb = 0.12593918054483;
        put(0, 1.0 / 3, 1.0 / 3, 0.225);
        put(1, a1, a2, wa); put(2, a2, a1, wa); put(3, a2, a2, wa);
        put(4, b1, b2, wb); put(5, b2, b1, wb); put(6, b2, b2, wb);
    }
    return q;
}
