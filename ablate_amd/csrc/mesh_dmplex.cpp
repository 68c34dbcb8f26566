// Mesh ingestion from a DMPlex topology: the cone of every point (DMPlexGetCone), the vertex coordinates and the
// "Face Sets" label, as an ABLATE domain holds them after DMPlexInterpolate and as DMPlexView_HDF5 stores them in
// the Hdf5Serializer output (src/io/hdf5Serializer.cpp:23-160: /topology/cones, /topology/cells,
// /geometry/vertices, /labels/Face Sets/<value>/indices).  It replaces, for such a domain, GhostBoundaryCells and
// DMPlexComputeGeometryFVM (src/domain/modifiers/ghostBoundaryCells.cpp:4-8, src/solver/cellSolver.cpp:193-196).
//
// Points of depth dim are cells, of depth dim-1 faces, of depth 0 vertices (depth = 1 + the largest depth in the
// cone).  Polygons (2-D cells, 3-D faces) are rebuilt as vertex loops by chaining their edges, so the orientation
// DMPlex keeps alongside the cones is not needed; a polyhedron's faces are oriented away from its vertex mean
// (convex cells).  Cells whose cone holds a single face are the ghost cells of DMPlexConstructGhostCells: they are
// dropped and made again (one per boundary face, interior centroid reflected across the face).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fv_internal.hpp"
#include "fv_layout.hpp"

using abl::HostMesh;

extern "C" int abl_fv_mesh_dmplex(int32_t dim, int64_t n_points, const int64_t* cone_offsets, const int64_t* cones,
                                  const double* vertex_coords, const int32_t* face_sets, int32_t ordering,
                                  abl_fv_mesh_handle* out) {
    using namespace abl;
    const std::string who = "abl_fv_mesh_dmplex";
    try {
        if (!cone_offsets || !cones || !vertex_coords || !out) throw std::invalid_argument(who + ": null argument");
        *out = nullptr;
        if (dim != 2 && dim != 3) throw std::invalid_argument(who + ": dim must be 2 or 3");
        if (n_points <= 0) throw std::invalid_argument(who + ": no points");
        const int64_t n = n_points;
        const int64_t* co = cone_offsets;
        if (co[0] != 0) throw std::invalid_argument(who + ": cone_offsets[0] must be 0");
        for (int64_t p = 0; p < n; p++)
            if (co[p + 1] < co[p]) throw std::invalid_argument(who + ": cone_offsets must not decrease");
        for (int64_t q = 0; q < co[n]; q++)
            if (cones[q] < 0 || cones[q] >= n) throw std::invalid_argument(who + ": a cone names a point out of range");

        // depth of every point: 0 for an empty cone, else 1 + the largest depth in the cone
        std::vector<int> depth(n, -1);
        for (int pass = 0; pass <= dim + 1; pass++)
            for (int64_t p = 0; p < n; p++) {
                if (depth[p] >= 0) continue;
                int mx = -1;
                bool known = true;
                for (int64_t q = co[p]; q < co[p + 1] && known; q++) {
                    known = depth[cones[q]] >= 0;
                    mx = std::max(mx, depth[cones[q]]);
                }
                if (known) depth[p] = mx + 1;
            }
        for (int64_t p = 0; p < n; p++) {
            if (depth[p] < 0 || depth[p] > dim)
                throw std::invalid_argument(who + ": point " + std::to_string(p) + " is deeper than the mesh dimension");
        }
        std::vector<int64_t> vid(n, -1);
        int64_t nv = 0;
        for (int64_t p = 0; p < n; p++)
            if (depth[p] == 0) vid[p] = nv++;
        std::vector<double> X(3 * nv, 0.0);
        for (int64_t k = 0; k < nv; k++)
            for (int d = 0; d < dim; d++) X[3 * k + d] = vertex_coords[(int64_t)dim * k + d];

        auto edge = [&](int64_t e) {
            if (depth[e] != 1 || co[e + 1] - co[e] != 2)
                throw std::invalid_argument(who + ": point " + std::to_string(e) + " is not an edge of two vertices");
            return std::make_pair(vid[cones[co[e]]], vid[cones[co[e] + 1]]);
        };
        // vertex loop of a polygon point (a 2-D cell, a 3-D face) from the edges of its cone
        std::vector<std::pair<int64_t, int64_t>> es;
        std::vector<char> used;
        auto loop = [&](int64_t p, std::vector<int64_t>& v) {
            const int64_t k = co[p + 1] - co[p];
            if (k < 3) throw std::invalid_argument(who + ": polygon " + std::to_string(p) + " has fewer than 3 edges");
            es.resize(k);
            used.assign(k, 0);
            for (int64_t i = 0; i < k; i++) es[i] = edge(cones[co[p] + i]);
            v.assign(1, es[0].first);
            int64_t cur = es[0].second;
            used[0] = 1;
            for (int64_t step = 1; step < k; step++) {
                v.push_back(cur);
                int64_t next = -1;
                for (int64_t i = 0; i < k && next < 0; i++)
                    if (!used[i] && (es[i].first == cur || es[i].second == cur)) {
                        used[i] = 1;
                        next = es[i].first == cur ? es[i].second : es[i].first;
                    }
                if (next < 0) throw std::invalid_argument(who + ": the edges of polygon " + std::to_string(p) + " do not form a loop");
                cur = next;
            }
            if (cur != v[0]) throw std::invalid_argument(who + ": the edges of polygon " + std::to_string(p) + " do not close");
        };

        // cells and the supports of their faces
        std::vector<int64_t> cells;
        for (int64_t p = 0; p < n; p++) {
            if (depth[p] != dim) continue;
            const int64_t k = co[p + 1] - co[p];
            if (k == 1) continue;  // ghost cell of DMPlexConstructGhostCells
            if (k < dim + 1) throw std::invalid_argument(who + ": cell " + std::to_string(p) + " has too few faces");
            cells.push_back(p);
        }
        const int64_t NI = (int64_t)cells.size();
        if (NI == 0) throw std::invalid_argument(who + ": no cells of depth dim (the plex must be interpolated)");
        if (NI >= ((int64_t)1 << 31)) throw std::invalid_argument(who + ": too many cells for 32-bit indices");
        std::vector<int32_t> s0(n, -1), s1(n, -1);
        for (int64_t c = 0; c < NI; c++)
            for (int64_t q = co[cells[c]]; q < co[cells[c] + 1]; q++) {
                const int64_t f = cones[q];
                if (depth[f] != dim - 1) throw std::invalid_argument(who + ": a cell cone holds a point that is not a face");
                if (s0[f] < 0) s0[f] = (int32_t)c;
                else if (s1[f] < 0) s1[f] = (int32_t)c;
                else throw std::invalid_argument(who + ": face " + std::to_string(f) + " supports more than two cells");
            }

        // cell geometry (point order)
        std::vector<double> vol(NI), cent(3 * NI, 0.0);
        std::vector<int64_t> source(NI);
        std::vector<int64_t> lv, cv;
        std::vector<const double*> v;
        for (int64_t c = 0; c < NI; c++) {
            const int64_t p = cells[c];
            source[c] = p;
            double* xc = &cent[3 * c];
            if (dim == 2) {
                loop(p, lv);
                v.clear();
                for (auto k : lv) v.push_back(&X[3 * k]);
                polygon_geometry(v, vol[c], xc);
            } else {
                // tetrahedral fan about the vertex mean m, each face oriented away from m
                cv.clear();
                for (int64_t q = co[p]; q < co[p + 1]; q++) {
                    loop(cones[q], lv);
                    cv.insert(cv.end(), lv.begin(), lv.end());
                }
                std::sort(cv.begin(), cv.end());
                cv.erase(std::unique(cv.begin(), cv.end()), cv.end());
                double m[3] = {0, 0, 0};
                for (auto k : cv)
                    for (int d = 0; d < 3; d++) m[d] += X[3 * k + d] / (double)cv.size();
                double V = 0.0, cs[3] = {0, 0, 0};
                for (int64_t q = co[p]; q < co[p + 1]; q++) {
                    loop(cones[q], lv);
                    v.clear();
                    for (auto k : lv) v.push_back(&X[3 * k]);
                    double nf[3], xf[3];
                    face_geometry3(v, nf, xf);
                    const double sgn =
                        nf[0] * (xf[0] - m[0]) + nf[1] * (xf[1] - m[1]) + nf[2] * (xf[2] - m[2]) < 0 ? -1.0 : 1.0;
                    for (size_t i = 1; i + 1 < v.size(); i++) {
                        double a[3], b[3], e[3];
                        for (int d = 0; d < 3; d++) {
                            a[d] = v[0][d] - m[d];
                            b[d] = v[i][d] - m[d];
                            e[d] = v[i + 1][d] - m[d];
                        }
                        const double t = sgn *
                                         (a[0] * (b[1] * e[2] - b[2] * e[1]) - a[1] * (b[0] * e[2] - b[2] * e[0]) +
                                          a[2] * (b[0] * e[1] - b[1] * e[0])) / 6.0;
                        V += t;
                        for (int d = 0; d < 3; d++) cs[d] += t * (m[d] + v[0][d] + v[i][d] + v[i + 1][d]) / 4.0;
                    }
                }
                vol[c] = V;
                for (int d = 0; d < 3; d++) xc[d] = cs[d] / V;
            }
            if (!(vol[c] > 0.0)) throw std::invalid_argument(who + ": degenerate cell " + std::to_string(p));
        }

        // faces (point order): the edge's two vertices in 2-D, the face loop in 3-D
        PolyFaces pf;
        for (int64_t f = 0; f < n; f++) {
            if (depth[f] != dim - 1 || s0[f] < 0) continue;
            int32_t slot = -1;
            if (s1[f] < 0) {
                slot = face_sets ? face_sets[f] : -1;
                if (slot < 0)
                    throw std::invalid_argument(who + ": boundary face " + std::to_string(f) + " has no \"Face Sets\" value");
            }
            if (dim == 2) {
                const auto uv = edge(f);
                const int64_t ev[2] = {uv.first, uv.second};
                pf.add(s0[f], s1[f], slot, ev, 2);
            } else {
                loop(f, lv);
                pf.add(s0[f], s1[f], slot, lv.data(), (int)lv.size());
            }
        }
        auto* mesh = new abl_fv_mesh();
        try {
            assemble_polytope_mesh(dim, X, vol, cent, pf, ordering, source, "abl_fv_mesh_dmplex", mesh->m);
        } catch (...) {
            delete mesh;
            throw;
        }
        *out = mesh;
        return 0;
    } catch (const std::exception& e) {
        abl::set_global_error(e.what());
        return 1;
    }
}
