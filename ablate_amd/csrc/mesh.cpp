// Mesh generators and orderings for the finite-volume RHS (host side).
//
// They produce the DMPlex finite-volume geometry the reference obtains from BoxMesh + GhostBoundaryCells +
// DMPlexComputeGeometryFVM (src/domain/boxMesh.cpp:17-43, src/domain/modifiers/ghostBoundaryCells.cpp:4-8,
// src/solver/cellSolver.cpp:193-196):
//   - face area-normals point from support[0] (left) to support[1] (right) with |n| = area;
//   - one ghost cell per boundary face, numbered after the interior cells, with the interior centroid reflected
//     across the face plane and the interior volume (the DMPlex ghost-cell convention);
//   - periodic directions produce wrap faces whose right cell is translated into the left cell's frame by
//     face_shift (the localized-coordinate image).
// Cell orderings: natural (x fastest), brick (8^3 / 16^2 bricks in Morton order of the bricks; a 3-D brick walked as
// four 8x4x4 blocks), column (3-D boxes: 8 x 8 columns walked layer by layer in z), Morton order of the cell
// centroids (unstructured meshes) and reverse Cuthill-McKee.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <stdexcept>

#include "fv_internal.hpp"

namespace abl {

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform in [-1, 1) from a counter
inline double uniform_pm1(uint64_t seed, uint64_t counter) {
    uint64_t r = splitmix64(seed ^ splitmix64(counter));
    return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t code = 0;
    for (int b = 0; b < 21; b++) {
        code |= (uint64_t)((x >> b) & 1u) << (3 * b);
        code |= (uint64_t)((y >> b) & 1u) << (3 * b + 1);
        code |= (uint64_t)((z >> b) & 1u) << (3 * b + 2);
    }
    return code;
}

// Order of the interior cells of a box: returns logical index (x fastest) for each new index.
std::vector<int64_t> box_order(int dim, const int32_t* n, int32_t ordering) {
    int64_t N = 1;
    for (int d = 0; d < dim; d++) N *= n[d];
    std::vector<int64_t> order(N);
    if (ordering == ABL_FV_ORDER_NATURAL) {
        std::iota(order.begin(), order.end(), 0);
        return order;
    }
    if (ordering == ABL_FV_ORDER_BRICK) {
        const int B = dim == 3 ? 8 : 16;
        int64_t nb[3] = {1, 1, 1};
        for (int d = 0; d < dim; d++) nb[d] = (n[d] + B - 1) / B;
        std::vector<std::pair<uint64_t, int64_t>> bricks;
        for (int64_t bz = 0; bz < nb[2]; bz++)
            for (int64_t by = 0; by < nb[1]; by++)
                for (int64_t bx = 0; bx < nb[0]; bx++)
                    bricks.push_back({morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz), bx + nb[0] * (by + nb[1] * bz)});
        std::sort(bricks.begin(), bricks.end());
        int64_t k = 0;
        const int64_t nz = dim == 3 ? n[2] : 1;
        // 3-D bricks are walked as four 8x4x4 blocks (y halves inside z halves), so that every 128 consecutive cells
        // of a whole brick form an 8x4x4 block and every 256 an 8x8x4 slab (compact cell tiles)
        const int SB = dim == 3 ? B / 2 : B;
        for (auto& br : bricks) {
            int64_t b = br.second;
            int64_t bx = b % nb[0], by = (b / nb[0]) % nb[1], bz = b / (nb[0] * nb[1]);
            int64_t z0 = dim == 3 ? bz * B : 0, z1 = dim == 3 ? std::min<int64_t>(nz, z0 + B) : 1;
            int64_t y0 = by * B, y1 = std::min<int64_t>(n[1], y0 + B);
            for (int64_t zs = z0; zs < z1; zs += SB)
                for (int64_t ys = y0; ys < y1; ys += SB)
                    for (int64_t z = zs; z < std::min<int64_t>(z1, zs + SB); z++)
                        for (int64_t y = ys; y < std::min<int64_t>(y1, ys + SB); y++)
                            for (int64_t x = bx * B; x < std::min<int64_t>(n[0], bx * B + B); x++)
                                order[k++] = x + (int64_t)n[0] * (y + (int64_t)n[1] * z);
        }
        return order;
    }
    if (ordering == ABL_FV_ORDER_COLUMN) {
        // columns of B x B cells in (x, y), x-major; inside a column z-layers, each layer y-major x-minor
        const int B = 8;
        const int64_t ncx = (n[0] + B - 1) / B, ncy = (n[1] + B - 1) / B;
        int64_t k = 0;
        for (int64_t cy = 0; cy < ncy; cy++)
            for (int64_t cx = 0; cx < ncx; cx++)
                for (int64_t z = 0; z < n[2]; z++)
                    for (int64_t y = cy * B; y < std::min<int64_t>(n[1], cy * B + B); y++)
                        for (int64_t x = cx * B; x < std::min<int64_t>(n[0], cx * B + B); x++)
                            order[k++] = x + (int64_t)n[0] * (y + (int64_t)n[1] * z);
        return order;
    }
    // RCM over the box's face graph (periodicity ignored for the ordering)
    std::vector<int64_t> off(N + 1, 0);
    std::vector<int32_t> adj;
    adj.reserve((size_t)N * 2 * dim);
    for (int64_t c = 0; c < N; c++) {
        int64_t ijk[3] = {c % n[0], (c / n[0]) % n[1], dim == 3 ? c / ((int64_t)n[0] * n[1]) : 0};
        for (int d = 0; d < dim; d++)
            for (int s = -1; s <= 1; s += 2) {
                int64_t q[3] = {ijk[0], ijk[1], ijk[2]};
                q[d] += s;
                if (q[d] < 0 || q[d] >= n[d]) continue;
                adj.push_back((int32_t)(q[0] + (int64_t)n[0] * (q[1] + (int64_t)n[1] * q[2])));
            }
        off[c + 1] = (int64_t)adj.size();
    }
    return rcm_order(N, off, adj);
}

}  // namespace

void finish_ghost(HostMesh& m, int64_t f, int32_t interior, int32_t ghost) {
    // reflect the interior centroid across the face plane (DMPlexComputeGeometryFVM ghost convention)
    const double* n = &m.face_normal[3 * f];
    const double* xf = &m.face_centroid[3 * f];
    const double* xi = &m.cell_centroid[3 * interior];
    double c2f[3], nn = 0.0, dot = 0.0;
    for (int d = 0; d < 3; d++) {
        c2f[d] = xf[d] - xi[d];
        nn += n[d] * n[d];
        dot += c2f[d] * n[d];
    }
    double a = dot / nn;
    for (int d = 0; d < 3; d++) m.cell_centroid[3 * ghost + d] = xi[d] + 2.0 * a * n[d];
    m.cell_volume[ghost] = m.cell_volume[interior];
}

std::vector<int64_t> morton_order(const std::vector<double>& xc, int64_t n) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t c = 0; c < n; c++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], xc[3 * c + d]);
            hi[d] = std::max(hi[d], xc[3 * c + d]);
        }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<std::pair<uint64_t, int64_t>> key(n);
    for (int64_t c = 0; c < n; c++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) q[d] = (uint32_t)std::min(2097151.0, (xc[3 * c + d] - lo[d]) / span * 2097151.0);
        key[c] = {morton3(q[0], q[1], q[2]), c};
    }
    std::sort(key.begin(), key.end());
    std::vector<int64_t> order(n);
    for (int64_t c = 0; c < n; c++) order[c] = key[c].second;
    return order;
}

std::vector<int64_t> rcm_order(int64_t n, const std::vector<int64_t>& off, const std::vector<int32_t>& adj) {
    std::vector<int64_t> perm;
    perm.reserve(n);
    std::vector<char> seen(n, 0);
    auto degree = [&](int64_t v) { return off[v + 1] - off[v]; };
    auto bfs_last = [&](int64_t start) {
        // farthest vertex (min degree among the last level) from start, within its component
        std::vector<int64_t> level{start}, next;
        std::vector<char> vis(0);
        std::vector<int64_t> touched{start};
        seen[start] = 2;
        int64_t best = start;
        while (!level.empty()) {
            next.clear();
            for (int64_t v : level)
                for (int64_t k = off[v]; k < off[v + 1]; k++) {
                    int64_t w = adj[k];
                    if (seen[w] == 0) {
                        seen[w] = 2;
                        next.push_back(w);
                        touched.push_back(w);
                    }
                }
            if (next.empty()) {
                best = level[0];
                for (int64_t v : level)
                    if (degree(v) < degree(best)) best = v;
            }
            level.swap(next);
        }
        for (int64_t v : touched) seen[v] = 0;
        return best;
    };
    for (int64_t root = 0; root < n; root++) {
        if (seen[root]) continue;
        int64_t s = bfs_last(bfs_last(root));
        size_t begin = perm.size();
        perm.push_back(s);
        seen[s] = 1;
        std::vector<int64_t> nb;
        for (size_t head = begin; head < perm.size(); head++) {
            int64_t v = perm[head];
            nb.clear();
            for (int64_t k = off[v]; k < off[v + 1]; k++)
                if (!seen[adj[k]]) {
                    seen[adj[k]] = 1;
                    nb.push_back(adj[k]);
                }
            std::sort(nb.begin(), nb.end(), [&](int64_t a, int64_t b) {
                return degree(a) != degree(b) ? degree(a) < degree(b) : a < b;
            });
            for (int64_t w : nb) perm.push_back(w);
        }
    }
    std::reverse(perm.begin(), perm.end());
    return perm;
}

}  // namespace abl

using abl::HostMesh;

extern "C" int abl_fv_mesh_box(int32_t dim, const int32_t* cells, const double* lower, const double* upper,
                               const int32_t* periodic, int32_t ordering, abl_fv_mesh_handle* out) {
    try {
        if (!out || !cells || !lower || !upper) throw std::invalid_argument("abl_fv_mesh_box: null argument");
        if (dim != 2 && dim != 3) throw std::invalid_argument("abl_fv_mesh_box: dim must be 2 or 3");
        if (ordering != ABL_FV_ORDER_NATURAL && ordering != ABL_FV_ORDER_BRICK && ordering != ABL_FV_ORDER_RCM &&
            !(ordering == ABL_FV_ORDER_COLUMN && dim == 3))
            throw std::invalid_argument("abl_fv_mesh_box: ordering must be natural, brick, RCM or (3-D) column");
        int32_t n[3] = {1, 1, 1};
        double h[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, len[3] = {1, 1, 1};
        int per[3] = {0, 0, 0};
        int64_t NI = 1;
        for (int d = 0; d < dim; d++) {
            n[d] = cells[d];
            if (n[d] < 1) throw std::invalid_argument("abl_fv_mesh_box: cells must be >= 1");
            lo[d] = lower[d];
            len[d] = upper[d] - lower[d];
            if (!(len[d] > 0)) throw std::invalid_argument("abl_fv_mesh_box: upper must exceed lower");
            h[d] = len[d] / n[d];
            per[d] = periodic ? (periodic[d] != 0) : 0;
            NI *= n[d];
        }
        if (NI >= (int64_t)1 << 31) throw std::invalid_argument("abl_fv_mesh_box: too many cells for 32-bit indices");
        auto* mesh = new abl_fv_mesh();
        HostMesh& m = mesh->m;
        m.dim = dim;
        std::vector<int64_t> order = abl::box_order(dim, n, ordering);
        std::vector<int32_t> inv(NI);
        for (int64_t k = 0; k < NI; k++) inv[order[k]] = (int32_t)k;
        m.source = order;  // lexicographic index (x fastest) of each cell
        int64_t nghost = 0;
        for (int d = 0; d < dim; d++)
            if (!per[d]) nghost += 2 * NI / n[d];
        int64_t NF = 0;
        for (int d = 0; d < dim; d++) NF += per[d] ? NI : (NI / n[d]) * (n[d] + 1);
        m.n_interior = NI;
        m.n_cells = NI + nghost;
        m.n_faces = NF;
        m.face_cells.resize(2 * NF);
        m.face_normal.assign(3 * NF, 0.0);
        m.face_centroid.assign(3 * NF, 0.0);
        m.face_shift.assign(3 * NF, 0.0);
        m.face_bc.assign(NF, -1);
        m.cell_centroid.assign(3 * m.n_cells, 0.0);
        m.cell_volume.assign(m.n_cells, 0.0);
        m.has_bc = nghost > 0;
        double vol = 1.0, area[3];
        for (int d = 0; d < dim; d++) vol *= h[d];
        for (int d = 0; d < dim; d++) area[d] = vol / h[d];
        for (int64_t k = 0; k < NI; k++) {
            int64_t c = order[k];
            int64_t ijk[3] = {c % n[0], (c / n[0]) % n[1], dim == 3 ? c / ((int64_t)n[0] * n[1]) : 0};
            for (int d = 0; d < dim; d++) m.cell_centroid[3 * k + d] = lo[d] + (ijk[d] + 0.5) * h[d];
            m.cell_volume[k] = vol;
        }
        int64_t f = 0, g = NI;
        for (int64_t k = 0; k < NI; k++) {
            int64_t c = order[k];
            int64_t ijk[3] = {c % n[0], (c / n[0]) % n[1], dim == 3 ? c / ((int64_t)n[0] * n[1]) : 0};
            const double* xc = &m.cell_centroid[3 * k];
            for (int d = 0; d < dim; d++) {
                if (!per[d] && ijk[d] == 0) {  // lower boundary face + ghost
                    m.face_cells[2 * f] = (int32_t)k;
                    m.face_cells[2 * f + 1] = (int32_t)g;
                    for (int e = 0; e < 3; e++) m.face_centroid[3 * f + e] = xc[e];
                    m.face_centroid[3 * f + d] -= 0.5 * h[d];
                    m.face_normal[3 * f + d] = -area[d];
                    m.face_bc[f] = 2 * d;
                    abl::finish_ghost(m, f, (int32_t)k, (int32_t)g);
                    f++;
                    g++;
                }
                int64_t q[3] = {ijk[0], ijk[1], ijk[2]};
                if (ijk[d] + 1 < n[d] || per[d]) {  // +d interior (possibly wrapping) face
                    bool wrap = ijk[d] + 1 == n[d];
                    q[d] = wrap ? 0 : ijk[d] + 1;
                    int64_t nbr = q[0] + (int64_t)n[0] * (q[1] + (int64_t)n[1] * q[2]);
                    m.face_cells[2 * f] = (int32_t)k;
                    m.face_cells[2 * f + 1] = inv[nbr];
                    for (int e = 0; e < 3; e++) m.face_centroid[3 * f + e] = xc[e];
                    m.face_centroid[3 * f + d] += 0.5 * h[d];
                    m.face_normal[3 * f + d] = area[d];
                    if (wrap) {
                        m.face_shift[3 * f + d] = len[d];
                        m.has_shift = true;
                    }
                    f++;
                } else {  // upper boundary face + ghost
                    m.face_cells[2 * f] = (int32_t)k;
                    m.face_cells[2 * f + 1] = (int32_t)g;
                    for (int e = 0; e < 3; e++) m.face_centroid[3 * f + e] = xc[e];
                    m.face_centroid[3 * f + d] += 0.5 * h[d];
                    m.face_normal[3 * f + d] = area[d];
                    m.face_bc[f] = 2 * d + 1;
                    abl::finish_ghost(m, f, (int32_t)k, (int32_t)g);
                    f++;
                    g++;
                }
            }
        }
        if (f != NF || g != m.n_cells) {
            delete mesh;
            throw std::logic_error("abl_fv_mesh_box: internal face/ghost count mismatch");
        }
        if (!m.has_shift) m.face_shift.clear();
        if (!m.has_bc) m.face_bc.clear();
        *out = mesh;
        return 0;
    } catch (const std::exception& e) {
        abl::set_global_error(e.what());
        return 1;
    }
}

extern "C" int abl_fv_mesh_tet_box(int32_t n, double jitter, uint64_t seed, int32_t ordering, abl_fv_mesh_handle* out) {
    try {
        if (!out) throw std::invalid_argument("abl_fv_mesh_tet_box: null argument");
        if (n < 1 || n > 400) throw std::invalid_argument("abl_fv_mesh_tet_box: n must be in [1, 400]");
        if (jitter < 0 || jitter >= 0.5) throw std::invalid_argument("abl_fv_mesh_tet_box: jitter must be in [0, 0.5)");
        if (ordering != ABL_FV_ORDER_NATURAL && ordering != ABL_FV_ORDER_RCM && ordering != ABL_FV_ORDER_MORTON)
            throw std::invalid_argument("abl_fv_mesh_tet_box: ordering must be natural, RCM or Morton");
        const int64_t nn = n + 1;
        const double h = 1.0 / n;
        std::vector<double> X(3 * nn * nn * nn);
        for (int64_t k = 0; k < nn; k++)
            for (int64_t j = 0; j < nn; j++)
                for (int64_t i = 0; i < nn; i++) {
                    int64_t v = i + nn * (j + nn * k);
                    int64_t idx[3] = {i, j, k};
                    bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                    for (int d = 0; d < 3; d++) {
                        double x = idx[d] * h;
                        if (interior) x += jitter * h * abl::uniform_pm1(seed, (uint64_t)(3 * v + d));
                        X[3 * v + d] = x;
                    }
                }
        // Kuhn subdivision: 6 tets per cube sharing the (0,0,0)-(1,1,1) diagonal; consistent across cubes.
        static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        const int64_t NI = 6 * (int64_t)n * n * n;
        std::vector<int32_t> tet(4 * NI);
        for (int64_t k = 0; k < n; k++)
            for (int64_t j = 0; j < n; j++)
                for (int64_t i = 0; i < n; i++) {
                    int64_t cube = i + n * (j + n * k);
                    for (int p = 0; p < 6; p++) {
                        int64_t q[3] = {i, j, k};
                        int32_t* t = &tet[4 * (6 * cube + p)];
                        t[0] = (int32_t)(q[0] + nn * (q[1] + nn * q[2]));
                        for (int s = 0; s < 3; s++) {
                            q[perms[p][s]] += 1;
                            t[s + 1] = (int32_t)(q[0] + nn * (q[1] + nn * q[2]));
                        }
                    }
                }
        // faces: key = sorted node triple
        struct FaceRec {
            uint64_t key;
            int32_t tet;
            int32_t opp;  // local index of the opposite node
        };
        std::vector<FaceRec> recs(4 * NI);
        for (int64_t t = 0; t < NI; t++)
            for (int l = 0; l < 4; l++) {
                int32_t a[3], m = 0;
                for (int s = 0; s < 4; s++)
                    if (s != l) a[m++] = tet[4 * t + s];
                std::sort(a, a + 3);
                recs[4 * t + l] = {((uint64_t)a[0] << 42) | ((uint64_t)a[1] << 21) | (uint64_t)a[2], (int32_t)t, l};
            }
        std::sort(recs.begin(), recs.end(), [](const FaceRec& x, const FaceRec& y) {
            return x.key != y.key ? x.key < y.key : x.tet < y.tet;
        });
        // cell graph + face list (in old numbering)
        struct FacePair {
            int32_t a, b, node[3];  // b < 0: boundary
        };
        std::vector<FacePair> fp;
        fp.reserve(2 * NI + 6 * (int64_t)n * n * 2);
        for (size_t r = 0; r < recs.size();) {
            uint64_t key = recs[r].key;
            int32_t nodes[3] = {(int32_t)(key >> 42), (int32_t)((key >> 21) & 0x1FFFFF), (int32_t)(key & 0x1FFFFF)};
            if (r + 1 < recs.size() && recs[r + 1].key == key) {
                fp.push_back({recs[r].tet, recs[r + 1].tet, {nodes[0], nodes[1], nodes[2]}});
                r += 2;
            } else {
                fp.push_back({recs[r].tet, -1, {nodes[0], nodes[1], nodes[2]}});
                r += 1;
            }
        }
        recs.clear();
        recs.shrink_to_fit();
        // ordering of the tets
        std::vector<int64_t> order(NI);
        if (ordering == ABL_FV_ORDER_MORTON) {
            std::vector<double> xc(3 * NI);
            for (int64_t t = 0; t < NI; t++)
                for (int d = 0; d < 3; d++) {
                    double s = 0.0;
                    for (int q = 0; q < 4; q++) s += X[3 * tet[4 * t + q] + d];
                    xc[3 * t + d] = 0.25 * s;
                }
            order = abl::morton_order(xc, NI);
        } else if (ordering == ABL_FV_ORDER_RCM) {
            std::vector<int64_t> off(NI + 1, 0);
            for (auto& p : fp)
                if (p.b >= 0) {
                    off[p.a + 1]++;
                    off[p.b + 1]++;
                }
            for (int64_t c = 0; c < NI; c++) off[c + 1] += off[c];
            std::vector<int32_t> adj(off[NI]);
            std::vector<int64_t> fill(off.begin(), off.end() - 1);
            for (auto& p : fp)
                if (p.b >= 0) {
                    adj[fill[p.a]++] = p.b;
                    adj[fill[p.b]++] = p.a;
                }
            order = abl::rcm_order(NI, off, adj);
        } else {
            std::iota(order.begin(), order.end(), 0);
        }
        std::vector<int32_t> inv(NI);
        for (int64_t k = 0; k < NI; k++) inv[order[k]] = (int32_t)k;
        // faces numbered by (left cell, right cell) in the new numbering; boundary faces right = ghost
        struct FaceNew {
            int32_t L, R;  // R < 0 boundary
            int32_t idx;
        };
        std::vector<FaceNew> fn(fp.size());
        for (size_t i = 0; i < fp.size(); i++) {
            int32_t a = inv[fp[i].a];
            if (fp[i].b >= 0) {
                int32_t b = inv[fp[i].b];
                fn[i] = {std::min(a, b), std::max(a, b), (int32_t)i};
            } else {
                fn[i] = {a, -1, (int32_t)i};
            }
        }
        std::sort(fn.begin(), fn.end(), [](const FaceNew& x, const FaceNew& y) {
            if (x.L != y.L) return x.L < y.L;
            uint32_t rx = (uint32_t)x.R, ry = (uint32_t)y.R;  // boundary (R=-1) last
            return rx < ry;
        });
        auto* mesh = new abl_fv_mesh();
        HostMesh& m = mesh->m;
        m.dim = 3;
        m.source = order;  // index of each tetrahedron in generation order
        const int64_t NF = (int64_t)fn.size();
        int64_t nghost = 0;
        for (auto& x : fn) nghost += x.R < 0;
        m.n_interior = NI;
        m.n_cells = NI + nghost;
        m.n_faces = NF;
        m.face_cells.resize(2 * NF);
        m.face_normal.assign(3 * NF, 0.0);
        m.face_centroid.assign(3 * NF, 0.0);
        m.face_bc.assign(NF, -1);
        m.cell_centroid.assign(3 * m.n_cells, 0.0);
        m.cell_volume.assign(m.n_cells, 0.0);
        m.has_bc = nghost > 0;
        double minvol = 1e300;
        for (int64_t k = 0; k < NI; k++) {
            const int32_t* t = &tet[4 * order[k]];
            const double *a = &X[3 * t[0]], *b = &X[3 * t[1]], *c = &X[3 * t[2]], *d = &X[3 * t[3]];
            double u[3], v[3], w[3];
            for (int e = 0; e < 3; e++) {
                u[e] = b[e] - a[e];
                v[e] = c[e] - a[e];
                w[e] = d[e] - a[e];
                m.cell_centroid[3 * k + e] = 0.25 * (a[e] + b[e] + c[e] + d[e]);
            }
            double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
            m.cell_volume[k] = std::fabs(det) / 6.0;
            // the unjittered Kuhn tet has det = sign of its axis permutation; a sign change means inversion
            static const double parity[6] = {1, -1, -1, 1, 1, -1};
            minvol = std::min(minvol, parity[order[k] % 6] * det / 6.0);
        }
        if (!(minvol > 1e-6 * h * h * h)) {
            delete mesh;
            throw std::runtime_error("abl_fv_mesh_tet_box: degenerate tetrahedron (reduce jitter)");
        }
        int64_t g = NI;
        for (int64_t f = 0; f < NF; f++) {
            const FacePair& p = fp[fn[f].idx];
            const double *a = &X[3 * p.node[0]], *b = &X[3 * p.node[1]], *c = &X[3 * p.node[2]];
            double u[3], v[3], nrm[3], xf[3];
            for (int e = 0; e < 3; e++) {
                u[e] = b[e] - a[e];
                v[e] = c[e] - a[e];
                xf[e] = (a[e] + b[e] + c[e]) / 3.0;
            }
            nrm[0] = 0.5 * (u[1] * v[2] - u[2] * v[1]);
            nrm[1] = 0.5 * (u[2] * v[0] - u[0] * v[2]);
            nrm[2] = 0.5 * (u[0] * v[1] - u[1] * v[0]);
            int32_t L = fn[f].L;
            const double* xl = &m.cell_centroid[3 * L];
            double dir[3];
            int32_t R = fn[f].R;
            for (int e = 0; e < 3; e++) dir[e] = (R >= 0 ? m.cell_centroid[3 * R + e] : xf[e]) - xl[e];
            double s = nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2];
            double sign = s < 0 ? -1.0 : 1.0;
            for (int e = 0; e < 3; e++) {
                m.face_normal[3 * f + e] = sign * nrm[e];
                m.face_centroid[3 * f + e] = xf[e];
            }
            m.face_cells[2 * f] = L;
            if (R >= 0) {
                m.face_cells[2 * f + 1] = R;
            } else {
                m.face_cells[2 * f + 1] = (int32_t)g;
                // boundary slot: the cube side all three nodes lie on (0..5 = x-, x+, y-, y+, z-, z+)
                int slot = -1;
                for (int d = 0; d < 3 && slot < 0; d++) {
                    int64_t i0[3];
                    for (int q = 0; q < 3; q++) {
                        int64_t v0 = p.node[q];
                        int64_t idx[3] = {v0 % nn, (v0 / nn) % nn, v0 / (nn * nn)};
                        i0[q] = idx[d];
                    }
                    if (i0[0] == 0 && i0[1] == 0 && i0[2] == 0) slot = 2 * d;
                    else if (i0[0] == n && i0[1] == n && i0[2] == n) slot = 2 * d + 1;
                }
                m.face_bc[f] = slot;
                abl::finish_ghost(m, f, L, (int32_t)g);
                g++;
            }
        }
        *out = mesh;
        return 0;
    } catch (const std::exception& e) {
        abl::set_global_error(e.what());
        return 1;
    }
}

extern "C" int abl_fv_mesh_describe(abl_fv_mesh_handle mesh, abl_fv_mesh_desc* desc) {
    if (!mesh || !desc) {
        abl::set_global_error("abl_fv_mesh_describe: null argument");
        return 1;
    }
    const HostMesh& m = mesh->m;
    desc->dim = m.dim;
    desc->n_cells = m.n_cells;
    desc->n_interior = m.n_interior;
    desc->n_faces = m.n_faces;
    desc->face_cells = m.face_cells.data();
    desc->face_normal = m.face_normal.data();
    desc->face_centroid = m.face_centroid.data();
    desc->face_shift = m.face_shift.empty() ? nullptr : m.face_shift.data();
    desc->cell_centroid = m.cell_centroid.data();
    desc->cell_volume = m.cell_volume.data();
    desc->face_bc = m.face_bc.empty() ? nullptr : m.face_bc.data();
    return 0;
}

extern "C" int abl_fv_mesh_source_cells(abl_fv_mesh_handle mesh, int64_t* source) {
    if (!mesh || !source) {
        abl::set_global_error("abl_fv_mesh_source_cells: null argument");
        return 1;
    }
    const HostMesh& m = mesh->m;
    if ((int64_t)m.source.size() != m.n_interior) {
        abl::set_global_error("abl_fv_mesh_source_cells: the mesh records no source numbering");
        return 1;
    }
    std::copy(m.source.begin(), m.source.end(), source);
    return 0;
}

extern "C" void abl_fv_mesh_free(abl_fv_mesh_handle mesh) { delete mesh; }
