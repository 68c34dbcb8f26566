// Host-side construction of the device mesh layout for the FV RHS.
//
// Replaces the one-time work the reference does lazily on its first RHS call: the cell least-squares gradient
// weights of CellInterpolant (BuildGradientReconstruction_Internal, cellInterpolant.cpp:641-715, which stores them
// per face side in PetscFVFaceGeom::grad[2][3]) and the LeastSquaresAverage face stencils of FaceInterpolant
// (stencils/leastSquaresAverage.cpp:5-155), restated as per-face-side weights of an unlimited cell gradient.
//
// Layout (structure of arrays; every array is indexed by face or cell so that consecutive GPU lanes touch
// consecutive addresses):
//   ell[k * NI + c]        k-th face of interior cell c, encoded (face << 1) | side (side 1 = c is the right cell),
//                          faces sorted by number (the reference's accumulation order), -1 padding
//   fcell[s * F + f]       support cells
//   fna[4 * f + {0,1,2}]   unit normal (left -> right), fna[4 * f + 3] = |n|   (one 32-byte record per face)
//   fdx[2*dim*f + s*dim + d]  x_face - x_cell(side s) in that cell's frame    (one record per face)
//   fw, fwa  [2*dim*f + s*dim + d]  least-squares weight of cell `side` for the neighbour across f (cell / aux
//                          stencils)
//   inv_vol[c]
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "fv_layout.hpp"

namespace abl {

namespace {

// cyclic Jacobi eigen-decomposition of a symmetric n x n (n <= 3) matrix
void sym_eigen(int n, double A[9], double V[9], double w[3]) {
    for (int i = 0; i < 9; i++) V[i] = 0.0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double tau = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                double t = std::copysign(1.0, tau) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    double x = A[k * n + p], y = A[k * n + q];
                    A[k * n + p] = cs * x - sn * y;
                    A[k * n + q] = sn * x + cs * y;
                }
                for (int k = 0; k < n; k++) {
                    double x = A[p * n + k], y = A[q * n + k];
                    A[p * n + k] = cs * x - sn * y;
                    A[q * n + k] = sn * x + cs * y;
                }
                for (int k = 0; k < n; k++) {
                    double x = V[k * n + p], y = V[k * n + q];
                    V[k * n + p] = cs * x - sn * y;
                    V[k * n + q] = sn * x + cs * y;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i * n + i];
}

}  // namespace

// Minimum-norm least-squares weights (the pseudo-inverse PetscFVComputeGradient returns): full-rank stencils
// use the normal equations with an explicit inverse, rank-deficient ones an eigen pseudo-inverse.
void lsq_pinv_weights(int n, int dim, const double* dx, double* w) {
    double M[9] = {0};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++)
            for (int b = 0; b < dim; b++) M[a * dim + b] += dx[i * dim + a] * dx[i * dim + b];
    double Minv[9] = {0};
    double tr = 0.0;
    for (int a = 0; a < dim; a++) tr += M[a * dim + a];
    bool full = false;
    if (dim == 2) {
        double det = M[0] * M[3] - M[1] * M[2];
        if (det > 1e-10 * tr * tr) {
            Minv[0] = M[3] / det;
            Minv[1] = -M[1] / det;
            Minv[2] = -M[2] / det;
            Minv[3] = M[0] / det;
            full = true;
        }
    } else {
        double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
        double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
        if (det > 1e-10 * tr * tr * tr) {
            Minv[0] = c00 / det;
            Minv[1] = (M[2] * M[7] - M[1] * M[8]) / det;
            Minv[2] = (M[1] * M[5] - M[2] * M[4]) / det;
            Minv[3] = c01 / det;
            Minv[4] = (M[0] * M[8] - M[2] * M[6]) / det;
            Minv[5] = (M[2] * M[3] - M[0] * M[5]) / det;
            Minv[6] = c02 / det;
            Minv[7] = (M[1] * M[6] - M[0] * M[7]) / det;
            Minv[8] = (M[0] * M[4] - M[1] * M[3]) / det;
            full = true;
        }
    }
    if (!full) {
        double A[9], V[9], lam[3];
        std::memcpy(A, M, sizeof(A));
        sym_eigen(dim, A, V, lam);
        double lmax = 0.0;
        for (int k = 0; k < dim; k++) lmax = std::max(lmax, std::fabs(lam[k]));
        // eigenvalues of dx^T dx are accurate to ~eps * lmax: singular values below 1e-6 s_max count as zero
        for (int k = 0; k < dim; k++) {
            if (!(lam[k] > 1e-12 * lmax) || lmax == 0.0) continue;
            for (int a = 0; a < dim; a++)
                for (int b = 0; b < dim; b++) Minv[a * dim + b] += V[a * dim + k] * V[b * dim + k] / lam[k];
        }
    }
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) {
            double s = 0.0;
            for (int b = 0; b < dim; b++) s += Minv[a * dim + b] * dx[i * dim + b];
            w[i * dim + a] = s;
        }
}

void build_layout(const abl_fv_mesh_desc& m, const abl_fv_physics_desc& p, Layout& L) {
    const int dim = m.dim;
    if (dim != 2 && dim != 3) throw std::invalid_argument("mesh dim must be 2 or 3");
    if (m.n_interior <= 0 || m.n_interior > m.n_cells) throw std::invalid_argument("invalid n_interior / n_cells");
    if (m.n_cells >= ((int64_t)1 << 31) || m.n_faces >= ((int64_t)1 << 30))
        throw std::invalid_argument("mesh too large for 32-bit cell/face indices");
    if (!m.face_cells || !m.face_normal || !m.face_centroid || !m.cell_centroid || !m.cell_volume)
        throw std::invalid_argument("mesh descriptor has a null required array");
    if (p.n_species < 0 || p.n_species > 64) throw std::invalid_argument("n_species must be in [0, 64]");
    const int64_t N = m.n_cells, NI = m.n_interior, F = m.n_faces;
    L.dim = dim;
    L.N = N;
    L.NI = NI;
    L.F = F;
    L.ns = p.n_species;
    L.nc = dim + 2 + p.n_species;

    // ---- validation + incidence
    std::vector<int32_t> deg(N, 0);
    for (int64_t f = 0; f < F; f++) {
        int32_t a = m.face_cells[2 * f], b = m.face_cells[2 * f + 1];
        if (a < 0 || a >= N || b < 0 || b >= N || a == b) throw std::invalid_argument("face_cells out of range at face " + std::to_string(f));
        if (a >= NI) throw std::invalid_argument("the left cell of every face must be interior (face " + std::to_string(f) + ")");
        deg[a]++;
        deg[b]++;
        bool bnd = b >= NI;
        if (bnd && !m.face_bc) throw std::invalid_argument("boundary face without face_bc slots");
        if (m.face_bc) {
            if (bnd != (m.face_bc[f] >= 0)) throw std::invalid_argument("face_bc must be >= 0 exactly on boundary faces");
            if (bnd && m.face_bc[f] >= kMaxBcSlots) throw std::invalid_argument("face_bc slot exceeds the supported number");
        }
        double nn = 0.0;
        for (int d = 0; d < dim; d++) nn += m.face_normal[3 * f + d] * m.face_normal[3 * f + d];
        if (!(nn > 0)) throw std::invalid_argument("zero-area face " + std::to_string(f));
    }
    int K = 0;
    for (int64_t c = 0; c < NI; c++) {
        if (deg[c] == 0) throw std::invalid_argument("isolated interior cell " + std::to_string(c));
        K = std::max(K, deg[c]);
    }
    for (int64_t c = NI; c < N; c++)
        if (deg[c] != 1) throw std::invalid_argument("ghost cell " + std::to_string(c) + " must have exactly one face");
    if (K > kMaxFacesPerCell) throw std::invalid_argument("cells with more than " + std::to_string(kMaxFacesPerCell) + " faces are not supported");
    for (int64_t c = 0; c < N; c++)
        if (!(m.cell_volume[c] > 0)) throw std::invalid_argument("non-positive cell volume at cell " + std::to_string(c));
    L.K = K;

    L.ell.assign((size_t)K * NI, -1);
    std::vector<int32_t> fill(NI, 0);
    for (int64_t f = 0; f < F; f++)  // increasing face number per cell
        for (int s = 0; s < 2; s++) {
            int32_t c = m.face_cells[2 * f + s];
            if (c >= NI) continue;
            L.ell[(size_t)fill[c]++ * NI + c] = (int32_t)((f << 1) | s);
        }

    // ---- face geometry
    L.fcell.resize(2 * F);
    L.fna.assign((size_t)4 * F, 0.0);
    L.fdx.resize((size_t)2 * dim * F);
    auto shift = [&](int64_t f, int d) { return m.face_shift ? m.face_shift[3 * f + d] : 0.0; };
    for (int64_t f = 0; f < F; f++) {
        int32_t a = m.face_cells[2 * f], b = m.face_cells[2 * f + 1];
        L.fcell[f] = a;
        L.fcell[F + f] = b;
        double mag = 0.0;
        for (int d = 0; d < dim; d++) mag += m.face_normal[3 * f + d] * m.face_normal[3 * f + d];
        mag = std::sqrt(mag);
        L.fna[4 * f + 3] = mag;
        for (int d = 0; d < dim; d++) {
            L.fna[4 * f + d] = m.face_normal[3 * f + d] / mag;
            L.fdx[(size_t)2 * dim * f + 0 * dim + d] = m.face_centroid[3 * f + d] - m.cell_centroid[3 * a + d];
            L.fdx[(size_t)2 * dim * f + 1 * dim + d] = m.face_centroid[3 * f + d] - (m.cell_centroid[3 * b + d] + shift(f, d));
        }
        if (b >= NI) L.bfaces.push_back({(int32_t)f, m.face_bc[f]});
    }
    L.inv_vol.resize(N);
    for (int64_t c = 0; c < N; c++) L.inv_vol[c] = 1.0 / m.cell_volume[c];

    // neighbour offset of cell c across face f in c's frame: x_n - x_c
    auto nbr_dx = [&](int64_t f, int s, double* out) {
        int32_t a = m.face_cells[2 * f], b = m.face_cells[2 * f + 1];
        for (int d = 0; d < dim; d++) {
            double xa = m.cell_centroid[3 * a + d], xb = m.cell_centroid[3 * b + d] + shift(f, d);
            out[d] = s == 0 ? xb - xa : xa - xb;
        }
    };

    // ---- cell least-squares weights (non-boundary faces only) and aux (face-stencil) weights (all faces)
    const bool recon = p.reconstruction == ABL_FV_RECON_LEASTSQUARES && p.flux_calculator != ABL_FV_FLUX_NONE;
    const bool diffusion = p.transport == ABL_FV_TRANSPORT_SUTHERLAND
                               ? (p.transport_disabled & (ABL_FV_TRANSPORT_CONDUCTIVITY | ABL_FV_TRANSPORT_VISCOSITY |
                                                          ABL_FV_TRANSPORT_DIFFUSIVITY)) !=
                                     (ABL_FV_TRANSPORT_CONDUCTIVITY | ABL_FV_TRANSPORT_VISCOSITY |
                                      ABL_FV_TRANSPORT_DIFFUSIVITY)
                               : (p.k != 0.0 || p.mu != 0.0 || p.diff != 0.0);
    const bool has_bnd = !L.bfaces.empty();
    L.has_boundary = has_bnd;
    if (recon || diffusion) {
        L.fw.assign((size_t)2 * dim * F, 0.0);
        if (has_bnd && diffusion) L.fwa.assign((size_t)2 * dim * F, 0.0);
        std::vector<double> dx, w;
        std::vector<int64_t> used;
        std::vector<int> sides;
        for (int pass = 0; pass < ((has_bnd && diffusion) ? 2 : 1); pass++) {
            std::vector<double>& dst = pass == 0 ? L.fw : L.fwa;
            for (int64_t c = 0; c < NI; c++) {
                dx.clear();
                used.clear();
                sides.clear();
                for (int k = 0; k < K; k++) {
                    int32_t e = L.ell[(size_t)k * NI + c];
                    if (e < 0) break;
                    int64_t f = e >> 1;
                    int s = e & 1;
                    bool bnd = L.fcell[F + f] >= NI;
                    if (pass == 0 && bnd) continue;  // CellInterpolant excludes boundary faces
                    double v[3];
                    nbr_dx(f, s, v);
                    for (int d = 0; d < dim; d++) dx.push_back(v[d]);
                    used.push_back(f);
                    sides.push_back(s);
                }
                if (used.empty()) continue;
                w.resize(dx.size());
                lsq_pinv_weights((int)used.size(), dim, dx.data(), w.data());
                for (size_t i = 0; i < used.size(); i++)
                    for (int d = 0; d < dim; d++) dst[(size_t)2 * dim * used[i] + sides[i] * dim + d] = w[i * dim + d];
            }
            if (pass == 1) {
                // ghost cells: one-neighbour stencil, minimum-norm solution dx / |dx|^2 (ghost is always side 1)
                for (auto& bf : L.bfaces) {
                    int64_t f = bf.face;
                    double v[3];
                    nbr_dx(f, 1, v);
                    double wv[3];
                    lsq_pinv_weights(1, dim, v, wv);
                    for (int d = 0; d < dim; d++) dst[(size_t)2 * dim * f + 1 * dim + d] = wv[d];
                }
            }
        }
    }
}

void build_tiles(const Layout& L, int KM, int T, Tiles& out) {
    const int64_t NI = L.NI;
    out.T = T;
    out.n = (NI + T - 1) / T;
    out.nf_max = 0;
    out.start.assign((size_t)out.n + 1, 0);
    out.faces.clear();
    out.faces.reserve((size_t)NI * 4);
    out.slot.assign((size_t)KM * NI, -1);
    // Faces of a tile in ascending face number: with faces numbered by their left cell, consecutive lanes read
    // consecutive face records (a measured alternative that listed each cell's owned faces round by round doubled
    // the kernel's fetched bytes by breaking that coalescing).
    std::vector<int32_t> list;
    for (int64_t t = 0; t < out.n; t++) {
        const int64_t c0 = t * T, c1 = std::min<int64_t>(c0 + T, NI);
        list.clear();
        for (int k = 0; k < L.K; k++)
            for (int64_t c = c0; c < c1; c++) {
                const int32_t e = L.ell[(size_t)k * NI + c];
                if (e >= 0) list.push_back(e >> 1);
            }
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        for (int k = 0; k < L.K; k++)
            for (int64_t c = c0; c < c1; c++) {
                const int32_t e = L.ell[(size_t)k * NI + c];
                if (e < 0) continue;
                const int32_t local = (int32_t)(std::lower_bound(list.begin(), list.end(), e >> 1) - list.begin());
                out.slot[(size_t)k * NI + c] = (local << 1) | (e & 1);
            }
        if (out.faces.size() + list.size() > (size_t)INT32_MAX)
            throw std::invalid_argument("mesh too large for 32-bit tile face lists");
        out.faces.insert(out.faces.end(), list.begin(), list.end());
        out.start[(size_t)t + 1] = (int32_t)out.faces.size();
        out.nf_max = std::max(out.nf_max, (int)list.size());
    }
}

}  // namespace abl

namespace abl {

void build_fused_tiles(const Layout& L, int KM, int T, FusedTiles& out) {
    const int64_t NI = L.NI, F = L.F;
    out.T = T;
    out.n = (NI + T - 1) / T;
    out.nfm = 0;
    out.nem = 0;
    out.fstart.assign((size_t)out.n + 1, 0);
    out.hstart.assign((size_t)out.n + 1, 0);
    out.ffaces.clear();
    out.fcode.clear();
    out.hcell.clear();
    out.hinfo.clear();
    out.ffaces.reserve((size_t)NI * 4);
    out.fcode.reserve((size_t)NI * 4);
    out.fslot.assign((size_t)KM * NI, -1);
    std::vector<int32_t> list, halo, entry;
    auto local_face = [&](int32_t f) {
        return (int32_t)(std::lower_bound(list.begin(), list.end(), f) - list.begin());
    };
    for (int64_t t = 0; t < out.n; t++) {
        const int64_t c0 = t * T, c1 = std::min<int64_t>(c0 + T, NI);
        auto own = [&](int64_t c) { return c >= c0 && c < c1; };
        list.clear();
        for (int k = 0; k < L.K; k++)
            for (int64_t c = c0; c < c1; c++) {
                const int32_t e = L.ell[(size_t)k * NI + c];
                if (e >= 0) list.push_back(e >> 1);
            }
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        // halo: interior cells across the tile's faces that are not tile cells
        halo.clear();
        for (int32_t f : list)
            for (int s = 0; s < 2; s++) {
                const int64_t x = L.fcell[(size_t)s * F + f];
                if (x < NI && !own(x)) halo.push_back((int32_t)x);
            }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        // one face entry per (halo cell, tile face of that cell), in halo order then slot order
        entry.assign(list.size(), -1);
        int32_t ne = 0;
        for (int32_t h : halo) {
            int32_t mask = 0;
            const int32_t first = ne;
            for (int k = 0; k < L.K; k++) {
                const int32_t e = L.ell[(size_t)k * NI + h];
                if (e < 0) continue;
                const int32_t j = local_face(e >> 1);
                if (j < (int32_t)list.size() && list[(size_t)j] == (e >> 1)) {
                    mask |= 1 << k;
                    entry[(size_t)j] = ne++;
                }
            }
            out.hcell.push_back(h);
            out.hinfo.push_back(mask | (first << 8));
        }
        if ((int64_t)T + ne >= 0xFFFF) throw std::invalid_argument("fused tiles: too many halo faces in a tile");
        for (size_t j = 0; j < list.size(); j++) {
            const int32_t f = list[j];
            const int64_t a = L.fcell[f], b = L.fcell[(size_t)F + f];
            const bool la = own(a), lb = b < NI && own(b);
            if ((!la || (b < NI && !lb)) && entry[j] < 0) throw std::logic_error("fused tiles: a halo side without entry");
            const int32_t lc = la ? (int32_t)(a - c0) : T + entry[j];
            const int32_t rc = b >= NI ? 0xFFFF : (lb ? (int32_t)(b - c0) : T + entry[j]);
            out.fcode.push_back((int32_t)((uint32_t)lc | ((uint32_t)rc << 16)));
        }
        for (int k = 0; k < L.K; k++)
            for (int64_t c = c0; c < c1; c++) {
                const int32_t e = L.ell[(size_t)k * NI + c];
                if (e < 0) continue;
                out.fslot[(size_t)k * NI + c] = (local_face(e >> 1) << 1) | (e & 1);
            }
        if (out.ffaces.size() + list.size() > (size_t)INT32_MAX || out.hcell.size() > (size_t)INT32_MAX)
            throw std::invalid_argument("mesh too large for 32-bit fused tile lists");
        out.ffaces.insert(out.ffaces.end(), list.begin(), list.end());
        out.fstart[(size_t)t + 1] = (int32_t)out.ffaces.size();
        out.hstart[(size_t)t + 1] = (int32_t)out.hcell.size();
        out.nfm = std::max(out.nfm, (int)list.size());
        out.nem = std::max(out.nem, (int)ne);
    }
}

void check_fused_tiles(const Layout& L, int KM, const FusedTiles& ft) {
    const int64_t NI = L.NI, F = L.F;
    const int T = ft.T;
    auto bad = [](const std::string& what) { throw std::logic_error("fused tiles: " + what); };
    if (T <= 0 || T > 256 || KM < L.K || KM > 8) bad("tile size or slot width out of range");
    if ((int64_t)ft.fstart.size() != ft.n + 1 || (int64_t)ft.hstart.size() != ft.n + 1 || ft.n * T < NI ||
        (ft.n - 1) * T >= NI)
        bad("tile count does not cover the interior cells");
    if (ft.fcode.size() != ft.ffaces.size() || ft.hinfo.size() != ft.hcell.size() ||
        ft.fslot.size() != (size_t)KM * (size_t)NI)
        bad("list sizes disagree");
    if (ft.fstart[0] != 0 || ft.hstart[0] != 0 || ft.fstart[ft.n] != (int32_t)ft.ffaces.size() ||
        ft.hstart[ft.n] != (int32_t)ft.hcell.size())
        bad("list offsets do not span the lists");
    std::vector<int32_t> side_code;
    for (int64_t t = 0; t < ft.n; t++) {
        const int64_t c0 = t * T;
        const int nown = (int)std::min<int64_t>(T, NI - c0);
        const int32_t f0 = ft.fstart[t], nf = ft.fstart[t + 1] - f0;
        const int32_t h0 = ft.hstart[t], nh = ft.hstart[t + 1] - h0;
        if (nf < 0 || nf > ft.nfm || nh < 0) bad("face or halo count of a tile out of range");
        // halo entries: consecutive per halo cell in halo order, each bit of the mask a slot of the cell
        int32_t ne = 0;
        for (int32_t i = 0; i < nh; i++) {
            const int64_t h = ft.hcell[(size_t)(h0 + i)];
            const int32_t info = ft.hinfo[(size_t)(h0 + i)];
            if (h < 0 || h >= NI || (h >= c0 && h < c0 + nown)) bad("halo cell out of range or inside its tile");
            const int32_t mask = info & 0xFF;
            if ((info >> 8) != ne || (mask >> L.K) != 0) bad("halo entry index or slot mask inconsistent");
            for (int k = 0; k < L.K; k++) {
                if (!((mask >> k) & 1)) continue;
                if (L.ell[(size_t)k * NI + h] < 0) bad("halo slot mask names an empty slot");
                ne++;
            }
        }
        if (ne > ft.nem || (int64_t)T + ne >= 0xFFFF) bad("halo face entries exceed the tile's LDS rows");
        // side codes of each tile face against the mesh
        for (int32_t j = 0; j < nf; j++) {
            const int64_t f = ft.ffaces[(size_t)(f0 + j)];
            if (f < 0 || f >= F || (j > 0 && ft.ffaces[(size_t)(f0 + j - 1)] >= f)) bad("tile faces not ascending");
            const uint32_t code = (uint32_t)ft.fcode[(size_t)(f0 + j)];
            const int lc = (int)(code & 0xFFFF), rc = (int)(code >> 16);
            const int64_t a = L.fcell[(size_t)f], b = L.fcell[(size_t)F + (size_t)f];
            auto check_side = [&](int sc, int64_t cell) {
                if (sc < nown) {
                    if (c0 + sc != cell) bad("tile-cell side code names another cell");
                } else if (sc < T || sc >= T + ne) {
                    bad("side code outside the tile cells and halo entries");
                }
            };
            check_side(lc, a);
            if (rc == 0xFFFF) {
                if (b < NI || lc >= nown) bad("ghost face without a tile cell on its left");
            } else {
                if (b >= NI) bad("ghost face with a non-ghost side code");
                check_side(rc, b);
            }
            if (lc >= T && rc >= T && rc != 0xFFFF) bad("tile face with no tile cell");
        }
        // every halo entry names a tile face of its halo cell on the side where that cell is
        int32_t e = 0;
        for (int32_t i = 0; i < nh; i++) {
            const int64_t h = ft.hcell[(size_t)(h0 + i)];
            const int32_t mask = ft.hinfo[(size_t)(h0 + i)] & 0xFF;
            for (int k = 0; k < L.K; k++) {
                if (!((mask >> k) & 1)) continue;
                const int32_t s = L.ell[(size_t)k * NI + h];
                const int32_t f = s >> 1;
                const auto it = std::lower_bound(ft.ffaces.begin() + f0, ft.ffaces.begin() + f0 + nf, f);
                if (it == ft.ffaces.begin() + f0 + nf || *it != f) bad("halo entry names a face outside its tile");
                const uint32_t code = (uint32_t)ft.fcode[(size_t)(it - ft.ffaces.begin())];
                const int sc = (s & 1) ? (int)(code >> 16) : (int)(code & 0xFFFF);
                if (sc != T + e) bad("halo entry and face side code disagree");
                e++;
            }
        }
        // slots of the tile cells name their tile-local faces and sides
        for (int k = 0; k < KM; k++)
            for (int64_t c = c0; c < c0 + nown; c++) {
                const int32_t s = k < L.K ? L.ell[(size_t)k * NI + c] : -1;
                const int32_t v = ft.fslot[(size_t)k * NI + c];
                if (s < 0) {
                    if (v != -1) bad("padding slot not marked");
                    continue;
                }
                if (v < 0 || (v >> 1) >= nf || ft.ffaces[(size_t)(f0 + (v >> 1))] != (s >> 1) || (v & 1) != (s & 1))
                    bad("tile-cell slot names the wrong face or side");
            }
    }
}

void build_march_plan(const Layout& L, int KM, int T, int max_steps, MarchPlan& out) {
    const int64_t NI = L.NI, F = L.F;
    const int64_t nt = (NI + T - 1) / T;
    if (nt >= INT32_MAX / 2) throw std::invalid_argument("march plan: too many tiles");
    auto tile_of = [&](int64_t c) { return (int64_t)(c / T); };
    auto tile_end = [&](int64_t t) { return std::min<int64_t>((t + 1) * T, NI); };
    // faces shared between each tile and the others
    std::vector<std::vector<std::pair<int32_t, int32_t>>> partners((size_t)nt);
    for (int64_t t = 0; t < nt; t++) {
        auto& pl = partners[(size_t)t];
        for (int64_t c = t * T; c < tile_end(t); c++)
            for (int k = 0; k < L.K; k++) {
                const int32_t e = L.ell[(size_t)k * NI + c];
                if (e < 0) continue;
                const int64_t o = L.fcell[(size_t)(1 - (e & 1)) * F + (e >> 1)];
                if (o >= NI || tile_of(o) == t) continue;
                const int32_t to = (int32_t)tile_of(o);
                auto it = std::find_if(pl.begin(), pl.end(), [&](const std::pair<int32_t, int32_t>& q) { return q.first == to; });
                if (it == pl.end())
                    pl.push_back({to, 1});
                else
                    it->second++;
            }
    }
    auto shared = [&](int64_t a, int64_t b) -> int32_t {
        if (a < 0 || b < 0) return 0;
        for (const auto& q : partners[(size_t)a])
            if (q.first == b) return q.second;
        return 0;
    };
    // the tile sharing the most faces with t (at least T / 4 of them), other than x1 and x2
    auto best_partner = [&](int64_t t, int64_t x1, int64_t x2) -> int32_t {
        int32_t best = -1, cnt = std::max(1, T / 4) - 1;
        for (const auto& q : partners[(size_t)t])
            if (q.first != x1 && q.first != x2 && q.second > cnt) {
                best = q.first;
                cnt = q.second;
            }
        return best;
    };
    // chains of consecutive face-sharing tiles, split into sequences of at most max_steps steps
    out = MarchPlan();
    out.T = T;
    out.seq.push_back(0);
    std::vector<std::pair<int64_t, int64_t>> seqs;  // [a, b) tiles
    for (int64_t a = 0; a < nt;) {
        int64_t b = a + 1;
        while (b < nt && shared(b - 1, b) > 0) b++;
        const int64_t len = b - a, parts = (len + max_steps - 1) / max_steps;
        for (int64_t q = 0; q < parts; q++) seqs.push_back({a + len * q / parts, a + len * (q + 1) / parts});
        a = b;
    }
    out.n_seq = (int64_t)seqs.size();
    out.gslot.assign((size_t)KM * NI, -1);
    out.fstart.push_back(0);
    out.hstart.push_back(0);
    out.pfstart.push_back(0);
    std::vector<int32_t> list, prev_list, halo;
    std::vector<int32_t> entry;
    auto local = [](const std::vector<int32_t>& v, int32_t f) {
        const auto it = std::lower_bound(v.begin(), v.end(), f);
        return (it != v.end() && *it == f) ? (int32_t)(it - v.begin()) : -1;
    };
    for (size_t w = 0; w < seqs.size(); w++) {
        const int64_t a = seqs[w].first, b = seqs[w].second;
        const bool chain_before = a > 0 && shared(a - 1, a) > 0;
        // the first step's next tile is a + 1 whenever it shares faces with a (inside the sequence or the chain's
        // next sequence), so the prologue tile must be another one
        const int64_t next0 = (b - a > 1 || (a + 1 < nt && shared(a, a + 1) > 0)) ? a + 1 : -1;
        const int64_t pro = chain_before ? a - 1 : best_partner(a, next0, -1);
        out.pro.push_back((int32_t)pro);
        // prologue faces: between the prologue tile and the first tile
        prev_list.clear();
        if (pro >= 0) {
            for (int64_t c = a * T; c < tile_end(a); c++)
                for (int k = 0; k < L.K; k++) {
                    const int32_t e = L.ell[(size_t)k * NI + c];
                    if (e < 0) continue;
                    const int64_t o = L.fcell[(size_t)(1 - (e & 1)) * F + (e >> 1)];
                    if (o < NI && tile_of(o) == pro) prev_list.push_back(e >> 1);
                }
            std::sort(prev_list.begin(), prev_list.end());
            prev_list.erase(std::unique(prev_list.begin(), prev_list.end()), prev_list.end());
            for (int32_t f : prev_list) {
                auto code = [&](int64_t x) { return tile_of(x) == a ? (int32_t)(x - a * T) : (int32_t)(T + x - pro * T); };
                out.pffaces.push_back(f);
                out.pfcode.push_back((int32_t)((uint32_t)code(L.fcell[(size_t)f]) | ((uint32_t)code(L.fcell[(size_t)F + f]) << 16)));
            }
        }
        out.pfstart.push_back((int32_t)out.pffaces.size());
        out.nfm = std::max(out.nfm, (int)prev_list.size());
        int64_t pv = pro;
        for (int64_t t = a; t < b; t++) {
            int64_t nx;
            if (t + 1 < b)
                nx = t + 1;
            else if (t + 1 < nt && shared(t, t + 1) > 0)
                nx = t + 1;  // the next sequence of the same chain
            else
                nx = best_partner(t, pv, -1);
            out.stile.push_back((int32_t)t);
            out.snext.push_back((int32_t)nx);
            // faces of the tile's cells not shared with the previous tile
            list.clear();
            for (int64_t c = t * T; c < tile_end(t); c++)
                for (int k = 0; k < L.K; k++) {
                    const int32_t e = L.ell[(size_t)k * NI + c];
                    if (e < 0) continue;
                    const int64_t o = L.fcell[(size_t)(1 - (e & 1)) * F + (e >> 1)];
                    if (o < NI && pv >= 0 && tile_of(o) == pv) continue;
                    list.push_back(e >> 1);
                }
            std::sort(list.begin(), list.end());
            list.erase(std::unique(list.begin(), list.end()), list.end());
            // halo cells: interior sides outside the cur and next tiles
            halo.clear();
            for (int32_t f : list)
                for (int s = 0; s < 2; s++) {
                    const int64_t x = L.fcell[(size_t)s * F + f];
                    if (x < NI && tile_of(x) != t && tile_of(x) != nx) halo.push_back((int32_t)x);
                }
            std::sort(halo.begin(), halo.end());
            halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
            entry.assign(list.size(), -1);
            int32_t ne = 0;
            for (int32_t h : halo) {
                int32_t mask = 0;
                const int32_t first = ne;
                for (int k = 0; k < L.K; k++) {
                    const int32_t e = L.ell[(size_t)k * NI + h];
                    if (e < 0) continue;
                    const int32_t j = local(list, e >> 1);
                    if (j < 0) continue;
                    mask |= 1 << k;
                    entry[(size_t)j] = ne++;
                }
                out.hcell.push_back(h);
                out.hinfo.push_back(mask | (first << 8));
            }
            if (2 * (int64_t)T + ne >= 0xFFFF) throw std::invalid_argument("march plan: too many halo faces in a step");
            for (size_t j = 0; j < list.size(); j++) {
                const int32_t f = list[j];
                auto code = [&](int64_t x) -> int32_t {
                    if (x >= NI) return 0xFFFF;
                    if (tile_of(x) == t) return (int32_t)(x - t * T);
                    if (tile_of(x) == nx) return (int32_t)(T + x - nx * T);
                    return 2 * T + entry[j];
                };
                out.ffaces.push_back(f);
                out.fcode.push_back((int32_t)((uint32_t)code(L.fcell[(size_t)f]) | ((uint32_t)code(L.fcell[(size_t)F + f]) << 16)));
            }
            // gather codes of the tile's cells
            for (int64_t c = t * T; c < tile_end(t); c++)
                for (int k = 0; k < L.K; k++) {
                    const int32_t e = L.ell[(size_t)k * NI + c];
                    if (e < 0) continue;
                    const int64_t o = L.fcell[(size_t)(1 - (e & 1)) * F + (e >> 1)];
                    const bool from_prev = o < NI && pv >= 0 && tile_of(o) == pv;
                    const int32_t j = local(from_prev ? prev_list : list, e >> 1);
                    if (j < 0) throw std::logic_error("march plan: a face of a tile cell is in no list");
                    out.gslot[(size_t)k * NI + c] = (j << 2) | (from_prev ? 2 : 0) | (e & 1);
                }
            if (out.ffaces.size() > (size_t)INT32_MAX / 2 || out.hcell.size() > (size_t)INT32_MAX / 2)
                throw std::invalid_argument("mesh too large for 32-bit march plan lists");
            out.fstart.push_back((int32_t)out.ffaces.size());
            out.hstart.push_back((int32_t)out.hcell.size());
            out.nfm = std::max(out.nfm, (int)list.size());
            out.nem = std::max(out.nem, (int)ne);
            prev_list.swap(list);
            pv = t;
        }
        out.seq.push_back((int32_t)out.stile.size());
    }
    out.n_steps = (int64_t)out.stile.size();
}

void check_march_plan(const Layout& L, int KM, const MarchPlan& p) {
    const int64_t NI = L.NI, F = L.F;
    const int T = p.T;
    auto bad = [](const std::string& what) { throw std::logic_error("march plan: " + what); };
    const int64_t nt = (NI + T - 1) / T;
    auto tile_of = [&](int64_t c) { return (int64_t)(c / T); };
    auto ncells = [&](int64_t t) { return (int)(std::min<int64_t>((t + 1) * T, NI) - t * T); };
    if (T <= 0 || T > 128 || KM < L.K || KM > 8) bad("tile size or slot width out of range");
    if ((int64_t)p.seq.size() != p.n_seq + 1 || (int64_t)p.pro.size() != p.n_seq ||
        (int64_t)p.pfstart.size() != p.n_seq + 1 || (int64_t)p.stile.size() != p.n_steps ||
        (int64_t)p.snext.size() != p.n_steps || (int64_t)p.fstart.size() != p.n_steps + 1 ||
        (int64_t)p.hstart.size() != p.n_steps + 1 || p.gslot.size() != (size_t)KM * (size_t)NI ||
        p.fcode.size() != p.ffaces.size() || p.pfcode.size() != p.pffaces.size() || p.hinfo.size() != p.hcell.size())
        bad("list sizes disagree");
    if (p.seq[0] != 0 || p.seq[p.n_seq] != p.n_steps || p.fstart[0] != 0 || p.fstart[p.n_steps] != (int32_t)p.ffaces.size() ||
        p.hstart[0] != 0 || p.hstart[p.n_steps] != (int32_t)p.hcell.size() || p.pfstart[0] != 0 ||
        p.pfstart[p.n_seq] != (int32_t)p.pffaces.size())
        bad("list offsets do not span the lists");
    std::vector<char> seen((size_t)nt, 0);
    // a face list entry's side code for cell x of role offset r (0: cur / first tile, T: next / prologue tile)
    auto side_code = [](uint32_t code, int side) { return side ? (int)(code >> 16) : (int)(code & 0xFFFF); };
    for (int64_t w = 0; w < p.n_seq; w++) {
        const int32_t s0 = p.seq[w], s1 = p.seq[w + 1];
        if (s1 <= s0) bad("empty sequence");
        const int64_t pro = p.pro[w];
        const int64_t first = p.stile[s0];
        if (pro >= nt || pro == first) bad("prologue tile out of range");
        // prologue faces
        for (int32_t j = p.pfstart[w]; j < p.pfstart[w + 1]; j++) {
            if (pro < 0) bad("prologue faces without a prologue tile");
            const int64_t f = p.pffaces[(size_t)j];
            if (f < 0 || f >= F || (j > p.pfstart[w] && p.pffaces[(size_t)j - 1] >= f)) bad("prologue faces not ascending");
            for (int s = 0; s < 2; s++) {
                const int64_t x = L.fcell[(size_t)s * F + (size_t)f];
                const int code = side_code((uint32_t)p.pfcode[(size_t)j], s);
                if (x >= NI) bad("prologue face with a ghost side");
                const int64_t want = tile_of(x) == first ? x - first * T : (tile_of(x) == pro ? T + x - pro * T : -1);
                if (want < 0 || code != want) bad("prologue side code names another cell");
            }
        }
        if (p.pfstart[w + 1] - p.pfstart[w] > p.nfm) bad("prologue face count exceeds the LDS rows");
        int64_t pv = pro;
        for (int32_t s = s0; s < s1; s++) {
            const int64_t t = p.stile[(size_t)s], nx = p.snext[(size_t)s];
            if (t < 0 || t >= nt || seen[(size_t)t]) bad("a tile is handled by no step or by two");
            seen[(size_t)t] = 1;
            if (s > s0 && t != p.snext[(size_t)s - 1]) bad("a step's tile is not the records its previous step computed");
            if (nx >= nt || nx == t || (nx >= 0 && nx == pv)) bad("next tile out of range or equal to the cur / previous tile");
            const int32_t f0 = p.fstart[s], nf = p.fstart[s + 1] - f0;
            const int32_t h0 = p.hstart[s], nh = p.hstart[s + 1] - h0;
            if (nf < 0 || nf > p.nfm || nh < 0) bad("face or halo count of a step out of range");
            int32_t ne = 0;
            for (int32_t i = 0; i < nh; i++) {
                const int64_t h = p.hcell[(size_t)(h0 + i)];
                const int32_t info = p.hinfo[(size_t)(h0 + i)];
                if (h < 0 || h >= NI || tile_of(h) == t || tile_of(h) == nx) bad("halo cell out of range or in the cur / next tile");
                const int32_t mask = info & 0xFF;
                if ((info >> 8) != ne || (mask >> L.K) != 0) bad("halo entry index or slot mask inconsistent");
                for (int k = 0; k < L.K; k++)
                    if ((mask >> k) & 1) {
                        const int32_t e = L.ell[(size_t)k * NI + h];
                        if (e < 0) bad("halo slot mask names an empty slot");
                        const int64_t f = e >> 1;
                        const auto it = std::lower_bound(p.ffaces.begin() + f0, p.ffaces.begin() + f0 + nf, (int32_t)f);
                        if (it == p.ffaces.begin() + f0 + nf || *it != f) bad("halo entry names a face outside its step");
                        if (side_code((uint32_t)p.fcode[(size_t)(it - p.ffaces.begin())], e & 1) != 2 * T + ne)
                            bad("halo entry and face side code disagree");
                        ne++;
                    }
            }
            if (ne > p.nem || 2 * T + ne >= 0xFFFF) bad("halo entries exceed the step's LDS rows");
            for (int32_t j = 0; j < nf; j++) {
                const int64_t f = p.ffaces[(size_t)(f0 + j)];
                if (f < 0 || f >= F || (j > 0 && p.ffaces[(size_t)(f0 + j - 1)] >= f)) bad("step faces not ascending");
                bool touches = false;
                for (int s2 = 0; s2 < 2; s2++) {
                    const int64_t x = L.fcell[(size_t)s2 * F + (size_t)f];
                    const int code = side_code((uint32_t)p.fcode[(size_t)(f0 + j)], s2);
                    if (x >= NI) {
                        if (s2 == 0 || code != 0xFFFF) bad("ghost side code wrong");
                        continue;
                    }
                    if (pv >= 0 && tile_of(x) == pv) bad("a step evaluates a face of the previous tile");
                    if (tile_of(x) == t) {
                        touches = true;
                        if (code != x - t * T) bad("cur side code names another cell");
                    } else if (nx >= 0 && tile_of(x) == nx) {
                        if (code != T + x - nx * T) bad("next side code names another cell");
                    } else if (code < 2 * T || code >= 2 * T + ne) {
                        bad("halo side code out of range");
                    }
                }
                if (!touches) bad("a step evaluates a face of no cur cell");
            }
            // gather codes
            for (int64_t c = t * T; c < t * T + ncells(t); c++)
                for (int k = 0; k < KM; k++) {
                    const int32_t e = k < L.K ? L.ell[(size_t)k * NI + c] : -1;
                    const int32_t g = p.gslot[(size_t)k * NI + c];
                    if (e < 0) {
                        if (g != -1) bad("padding slot not marked");
                        continue;
                    }
                    if (g < 0 || (g & 1) != (e & 1)) bad("gather side wrong");
                    const bool from_prev = (g >> 1) & 1;
                    const int32_t j = g >> 2;
                    int64_t fj;
                    int code;
                    if (!from_prev) {
                        if (j >= nf) bad("gather index out of the step's list");
                        fj = p.ffaces[(size_t)(f0 + j)];
                        code = side_code((uint32_t)p.fcode[(size_t)(f0 + j)], e & 1);
                        if (code != c - t * T) bad("gather names a face entry of another cell");
                    } else if (s == s0) {
                        if (pro < 0 || j >= p.pfstart[w + 1] - p.pfstart[w]) bad("gather index out of the prologue list");
                        fj = p.pffaces[(size_t)(p.pfstart[w] + j)];
                        code = side_code((uint32_t)p.pfcode[(size_t)(p.pfstart[w] + j)], e & 1);
                        if (code != c - t * T) bad("gather names a prologue entry of another cell");
                    } else {
                        const int32_t pf0 = p.fstart[s - 1], pnf = p.fstart[s] - pf0;
                        if (j >= pnf) bad("gather index out of the previous step's list");
                        fj = p.ffaces[(size_t)(pf0 + j)];
                        code = side_code((uint32_t)p.fcode[(size_t)(pf0 + j)], e & 1);
                        if (code != T + c - t * T) bad("gather names a previous-step entry of another cell");
                    }
                    if (fj != (e >> 1)) bad("gather names another face");
                }
            pv = t;
        }
    }
    for (int64_t t = 0; t < nt; t++)
        if (!seen[(size_t)t]) bad("a tile is handled by no step");
}

bool lsq_stencils_full_rank(const Layout& L) {
    const int dim = L.dim;
    const int64_t NI = L.NI, F = L.F;
    for (int64_t c = 0; c < NI; c++) {
        double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < L.K; k++) {
            const int32_t e = L.ell[(size_t)k * NI + c];
            if (e < 0) continue;
            const int64_t f = e >> 1;
            const int s = e & 1;
            if (L.fcell[(size_t)F + f] >= NI) continue;  // ghost neighbours are not in the cell stencil
            double v[3] = {0, 0, 0};
            for (int d = 0; d < dim; d++)
                v[d] = L.fdx[(size_t)2 * dim * f + (size_t)s * dim + d] - L.fdx[(size_t)2 * dim * f + (size_t)(1 - s) * dim + d];
            for (int r = 0; r < dim; r++)
                for (int q = 0; q < dim; q++) M[r][q] += v[r] * v[q];
        }
        // condition number estimate ||M||_F ||M^-1||_F with M^-1 = adj(M) / det(M): the normal equations lose
        // cond(M) ulps against the pseudo-inverse, so poorly conditioned stencils keep the precomputed weights
        double adj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, det;
        if (dim == 3) {
            adj[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
            adj[0][1] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
            adj[0][2] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
            adj[1][0] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
            adj[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
            adj[1][2] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
            adj[2][0] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
            adj[2][1] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
            adj[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
            det = M[0][0] * adj[0][0] + M[0][1] * adj[1][0] + M[0][2] * adj[2][0];
        } else {
            adj[0][0] = M[1][1];
            adj[0][1] = -M[0][1];
            adj[1][0] = -M[1][0];
            adj[1][1] = M[0][0];
            det = M[0][0] * M[1][1] - M[0][1] * M[1][0];
        }
        double nm = 0.0, na = 0.0;
        for (int r = 0; r < dim; r++)
            for (int q = 0; q < dim; q++) {
                nm += M[r][q] * M[r][q];
                na += adj[r][q] * adj[r][q];
            }
        if (!(det > 0.0) || !(std::sqrt(nm) * std::sqrt(na) <= 1e4 * det)) return false;
    }
    return true;
}

}  // namespace abl
