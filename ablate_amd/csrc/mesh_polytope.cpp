// Finite-volume geometry of meshes given as polytopes (cells as vertex loops / faces as vertex loops), shared by
// the mesh-file readers (Gmsh files, DMPlex topology).  It restates what DMPlexComputeGeometryFVM yields for
// straight-sided cells with planar faces (src/solver/cellSolver.cpp:193-196 calls it): polygon area and centroid,
// edge midpoint and edge normal of length |edge| in 2-D, Newell area vector and fan centroid of polygon faces in
// 3-D, normals pointing from support[0] to support[1], and the ghost cells GhostBoundaryCells adds
// (src/domain/modifiers/ghostBoundaryCells.cpp:4-8) with the interior centroid reflected across the face.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv_internal.hpp"
#include "fv_layout.hpp"

namespace abl {

void polygon_geometry(const std::vector<const double*>& v, double& area, double* c) {
    double A = 0.0, cx = 0.0, cy = 0.0;
    const double x0 = v[0][0], y0 = v[0][1];
    for (size_t i = 1; i + 1 < v.size(); i++) {
        const double ax = v[i][0] - x0, ay = v[i][1] - y0, bx = v[i + 1][0] - x0, by = v[i + 1][1] - y0;
        const double t = 0.5 * (ax * by - ay * bx);  // signed triangle (v0, vi, vi+1)
        A += t;
        cx += t * (ax + bx) / 3.0;
        cy += t * (ay + by) / 3.0;
    }
    area = std::fabs(A);
    c[0] = x0 + cx / A;
    c[1] = y0 + cy / A;
    c[2] = 0.0;
}

void face_geometry3(const std::vector<const double*>& v, double* n, double* c) {
    n[0] = n[1] = n[2] = 0.0;
    double A = 0.0, cs[3] = {0, 0, 0};
    for (size_t i = 0; i < v.size(); i++) {
        const double* a = v[i];
        const double* b = v[(i + 1) % v.size()];
        n[0] += 0.5 * (a[1] * b[2] - a[2] * b[1]);
        n[1] += 0.5 * (a[2] * b[0] - a[0] * b[2]);
        n[2] += 0.5 * (a[0] * b[1] - a[1] * b[0]);
    }
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (size_t i = 1; i + 1 < v.size(); i++) {
        double u[3], w[3], x[3];
        for (int d = 0; d < 3; d++) {
            u[d] = v[i][d] - v[0][d];
            w[d] = v[i + 1][d] - v[0][d];
        }
        x[0] = u[1] * w[2] - u[2] * w[1];
        x[1] = u[2] * w[0] - u[0] * w[2];
        x[2] = u[0] * w[1] - u[1] * w[0];
        const double t = 0.5 * (x[0] * n[0] + x[1] * n[1] + x[2] * n[2]) / nn;  // signed triangle area
        A += t;
        for (int d = 0; d < 3; d++) cs[d] += t * (v[0][d] + v[i][d] + v[i + 1][d]) / 3.0;
    }
    for (int d = 0; d < 3; d++) c[d] = cs[d] / A;
}

void assemble_polytope_mesh(int dim, const std::vector<double>& X, const std::vector<double>& vol,
                            const std::vector<double>& cent, const PolyFaces& pf, int ordering,
                            const std::vector<int64_t>& source, const char* who, HostMesh& m) {
    const std::string w(who);
    const int64_t NI = (int64_t)vol.size();
    if (NI == 0) throw std::runtime_error(w + ": the mesh has no cells");
    if (NI >= ((int64_t)1 << 31)) throw std::runtime_error(w + ": too many cells for 32-bit indices");
    const int64_t NF = (int64_t)pf.a.size();
    std::vector<int64_t> order(NI);
    std::iota(order.begin(), order.end(), 0);
    if (ordering == ABL_FV_ORDER_MORTON) {
        order = morton_order(cent, NI);
    } else if (ordering == ABL_FV_ORDER_RCM) {
        std::vector<int64_t> off(NI + 1, 0);
        for (int64_t i = 0; i < NF; i++)
            if (pf.b[i] >= 0) {
                off[pf.a[i] + 1]++;
                off[pf.b[i] + 1]++;
            }
        for (int64_t c = 0; c < NI; c++) off[c + 1] += off[c];
        std::vector<int32_t> adj(off[NI]);
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int64_t i = 0; i < NF; i++)
            if (pf.b[i] >= 0) {
                adj[fill[pf.a[i]]++] = pf.b[i];
                adj[fill[pf.b[i]]++] = pf.a[i];
            }
        order = rcm_order(NI, off, adj);
    } else if (ordering != ABL_FV_ORDER_NATURAL) {
        throw std::invalid_argument(w + ": ordering must be natural, RCM or Morton");
    }
    std::vector<int32_t> inv(NI);
    for (int64_t k = 0; k < NI; k++) inv[order[k]] = (int32_t)k;

    int64_t nghost = 0;
    for (int64_t i = 0; i < NF; i++) nghost += pf.b[i] < 0;
    m.dim = dim;
    m.n_interior = NI;
    m.n_cells = NI + nghost;
    m.n_faces = NF;
    m.cell_centroid.assign(3 * m.n_cells, 0.0);
    m.cell_volume.assign(m.n_cells, 0.0);
    m.source.resize(NI);
    for (int64_t k = 0; k < NI; k++) {
        const int64_t c = order[k];
        m.cell_volume[k] = vol[c];
        for (int d = 0; d < 3; d++) m.cell_centroid[3 * k + d] = cent[3 * c + d];
        m.source[k] = source[c];
    }
    struct FaceNew {
        int32_t L, R;  // R < 0: boundary
        int64_t idx;
    };
    std::vector<FaceNew> fn(NF);
    for (int64_t i = 0; i < NF; i++) {
        const int32_t a = inv[pf.a[i]];
        if (pf.b[i] >= 0) {
            const int32_t b = inv[pf.b[i]];
            fn[i] = {std::min(a, b), std::max(a, b), i};
        } else {
            fn[i] = {a, -1, i};
        }
    }
    std::sort(fn.begin(), fn.end(), [](const FaceNew& x, const FaceNew& y) {
        if (x.L != y.L) return x.L < y.L;
        return (uint32_t)x.R < (uint32_t)y.R;  // boundary (R = -1) last
    });
    m.face_cells.resize(2 * NF);
    m.face_normal.assign(3 * NF, 0.0);
    m.face_centroid.assign(3 * NF, 0.0);
    m.face_bc.assign(NF, -1);
    m.has_bc = nghost > 0;
    int64_t g = NI;
    std::vector<const double*> v;
    for (int64_t fi = 0; fi < NF; fi++) {
        const int64_t i = fn[fi].idx;
        v.clear();
        for (int64_t q = pf.off[i]; q < pf.off[i + 1]; q++) v.push_back(&X[3 * pf.vtx[q]]);
        double nrm[3] = {0, 0, 0}, xf[3] = {0, 0, 0};
        if (dim == 2) {
            for (int d = 0; d < 2; d++) xf[d] = 0.5 * (v[0][d] + v[1][d]);
            nrm[0] = v[1][1] - v[0][1];
            nrm[1] = -(v[1][0] - v[0][0]);
        } else {
            face_geometry3(v, nrm, xf);
        }
        const int32_t L = fn[fi].L, R = fn[fi].R;
        const double* xl = &m.cell_centroid[3 * L];
        double s = 0.0;
        for (int d = 0; d < 3; d++) s += nrm[d] * (xf[d] - xl[d]);  // outward from the left cell
        const double sign = s < 0 ? -1.0 : 1.0;
        for (int d = 0; d < 3; d++) {
            m.face_normal[3 * fi + d] = sign * nrm[d];
            m.face_centroid[3 * fi + d] = xf[d];
        }
        m.face_cells[2 * fi] = L;
        if (R >= 0) {
            m.face_cells[2 * fi + 1] = R;
            continue;
        }
        if (pf.bc[i] < 0) throw std::runtime_error(w + ": a boundary face carries no boundary label value");
        if (pf.bc[i] >= kMaxBcSlots)
            throw std::runtime_error(w + ": boundary label values must be below " + std::to_string(kMaxBcSlots));
        m.face_cells[2 * fi + 1] = (int32_t)g;
        m.face_bc[fi] = pf.bc[i];
        finish_ghost(m, fi, L, (int32_t)g);
        g++;
    }
}

}  // namespace abl
