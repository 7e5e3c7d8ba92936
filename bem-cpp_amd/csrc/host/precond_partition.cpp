// Block partition of the RWG unknowns for the block-Jacobi / restricted
// additive Schwarz preconditioner (precond.cu:407-605, do_block_jacobi
// steps 1-2c), restated for the dense path.
//
//  1. RWG centres: the mean of the 2 Nq quadrature points of its two halves
//     (points r = l0 v0 + l1 v1 + l2 v2, bem_fmm.cu:250-300); the element
//     extent is the largest side of their bounding box, averaged over RWGs.
//  2. Cells of side bb_max / (2.5 / radius_mult) (at least the mean extent)
//     anchored at the centres' bounding-box minimum; the RWGs of one cell
//     form a block (sorted indices).
//  2b. Blocks above max_block_rwg are bisected at the midpoint of the longest
//     side of their centres' bounding box, for at most 21 rounds.
//  2c. With overlap_layers > 0 each block is extended by every RWG of the
//     other blocks whose centre lies in the block's bounding box grown by
//     overlap_layers * 2 * mean extent (own RWGs first, then the overlap in
//     index order).
// The reference keeps its cells in an unordered_map, so its block order is
// the hash order; here blocks follow the cell index (x, then y, then z) and
// bisection replaces a block by its two halves in place.  The partition is
// the same.  max_block_rwg == 1 gives one block per RWG: the 2 x 2
// self-interaction preconditioner of precond.cu:976-1090.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "bem_host.h"

BlockPartition block_partition(const Mesh& mesh, const RWG& rwg, const TriQuad& quad, double radius_mult,
                               int max_block_rwg, int overlap_layers) {
    const int N = rwg.N, Nq = quad.npts;
    BlockPartition P;
    std::vector<double> centre(3 * (size_t)N);
    double avg_extent = 0.0;
    for (int m = 0; m < N; m++) {
        double sum[3] = {0, 0, 0}, lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
        for (int half = 0; half < 2; half++) {
            Vec3 v0, v1, v2;
            mesh.tri_verts(half == 0 ? rwg.tri_p[m] : rwg.tri_m[m], v0, v1, v2);
            for (int q = 0; q < Nq; q++) {
                const double l1 = quad.pts[q][0], l2 = quad.pts[q][1], l0 = 1.0 - l1 - l2;
                const Vec3 r = v0 * l0 + v1 * l1 + v2 * l2;
                const double x[3] = {r.x, r.y, r.z};
                for (int d = 0; d < 3; d++) {
                    sum[d] += x[d];
                    lo[d] = std::min(lo[d], x[d]);
                    hi[d] = std::max(hi[d], x[d]);
                }
            }
        }
        for (int d = 0; d < 3; d++) centre[3 * (size_t)m + d] = sum[d] / (2.0 * Nq);
        avg_extent += std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    }
    avg_extent /= std::max(N, 1);
    P.avg_extent = avg_extent;

    std::vector<std::vector<int>> blocks;
    if (max_block_rwg == 1) {
        for (int m = 0; m < N; m++) blocks.push_back({m});
    } else {
        double bmin[3] = {1e30, 1e30, 1e30}, bmax[3] = {-1e30, -1e30, -1e30};
        for (int m = 0; m < N; m++)
            for (int d = 0; d < 3; d++) {
                bmin[d] = std::min(bmin[d], centre[3 * (size_t)m + d]);
                bmax[d] = std::max(bmax[d], centre[3 * (size_t)m + d]);
            }
        const double span = std::max({bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]});
        double cell = span / (2.5 / radius_mult);
        if (cell < avg_extent) cell = avg_extent;
        if (cell < 1e-15) cell = 1e-15;
        const double inv = 1.0 / cell;
        std::map<std::tuple<int, int, int>, std::vector<int>> cells;
        for (int m = 0; m < N; m++) {
            const double* c = &centre[3 * (size_t)m];
            cells[std::make_tuple((int)std::floor((c[0] - bmin[0]) * inv), (int)std::floor((c[1] - bmin[1]) * inv),
                                  (int)std::floor((c[2] - bmin[2]) * inv))]
                .push_back(m);  // m ascending: members stay sorted
        }
        for (auto& kv : cells) blocks.push_back(std::move(kv.second));
        P.cell_size = cell;
        // bisection of oversized blocks
        for (int round = 0; round < 21; round++) {
            bool split = false;
            std::vector<std::vector<int>> next;
            next.reserve(blocks.size());
            for (auto& b : blocks) {
                if ((int)b.size() <= max_block_rwg) {
                    next.push_back(std::move(b));
                    continue;
                }
                split = true;
                double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
                for (int m : b)
                    for (int d = 0; d < 3; d++) {
                        lo[d] = std::min(lo[d], centre[3 * (size_t)m + d]);
                        hi[d] = std::max(hi[d], centre[3 * (size_t)m + d]);
                    }
                int ax = 0;
                for (int d = 1; d < 3; d++)
                    if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
                const double mid = 0.5 * (lo[ax] + hi[ax]);
                std::vector<int> left, right;
                for (int m : b) (centre[3 * (size_t)m + ax] < mid ? left : right).push_back(m);
                if (left.empty() || right.empty()) {
                    next.push_back(std::move(b));
                } else {
                    next.push_back(std::move(left));
                    next.push_back(std::move(right));
                }
            }
            blocks = std::move(next);
            if (!split) break;
        }
    }
    P.own = blocks;
    P.ext = blocks;
    if (overlap_layers > 0 && max_block_rwg != 1) {
        std::vector<int> owner(N);
        for (size_t b = 0; b < blocks.size(); b++)
            for (int m : blocks[b]) owner[m] = (int)b;
        const double grow = overlap_layers * avg_extent * 2.0;
        for (size_t b = 0; b < blocks.size(); b++) {
            double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
            for (int m : blocks[b])
                for (int d = 0; d < 3; d++) {
                    lo[d] = std::min(lo[d], centre[3 * (size_t)m + d]);
                    hi[d] = std::max(hi[d], centre[3 * (size_t)m + d]);
                }
            for (int d = 0; d < 3; d++) {
                lo[d] -= grow;
                hi[d] += grow;
            }
            for (int m = 0; m < N; m++) {
                if (owner[m] == (int)b) continue;
                const double* c = &centre[3 * (size_t)m];
                if (c[0] >= lo[0] && c[0] <= hi[0] && c[1] >= lo[1] && c[1] <= hi[1] && c[2] >= lo[2] &&
                    c[2] <= hi[2])
                    P.ext[b].push_back(m);  // m ascending: the overlap stays sorted
            }
        }
    }
    return P;
}
