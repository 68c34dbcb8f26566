// Gmsh mesh ingestion (MSH 4.1 ASCII) into the finite-volume geometry of the RHS.
//
// Replaces, for the reference's `!ablate::domain::MeshFile` inputs (src/domain/meshFile.cpp, PETSc's
// DMPlexCreateGmshFromFile) followed by GhostBoundaryCells and DMPlexComputeGeometryFVM
// (src/domain/modifiers/ghostBoundaryCells.cpp:4-8, src/solver/cellSolver.cpp:193-196):
//   - cells are the elements of the highest dimension (triangles / quadrilaterals in 2-D; tetrahedra, hexahedra,
//     prisms and pyramids in 3-D, first order), faces the edges / polygons they share;
//   - boundary faces take the physical tag of the lower-dimensional element lying on them (the "Face Sets" label
//     value a boundary condition's labelIds refer to) and one ghost cell each (interior centroid reflected across
//     the face, interior volume);
//   - geometry is exact for straight-sided cells with planar faces: polygon area / centroid, polyhedron volume and
//     centroid from a tetrahedral fan about the vertex mean, face normals from Newell's formula (area vector);
//     normals point from the left cell (support[0]) to the right one.
// Faces are numbered by their left cell (the lower cell number); cell order natural (file order), RCM or Morton
// (numbering and face geometry: assemble_polytope_mesh).
#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv_internal.hpp"
#include "fv_layout.hpp"

namespace abl {
namespace {

struct ElemType {
    int dim, nodes;
};
// Gmsh element types of first order (MSH format documentation, "elm-type"): 1 line, 2 triangle, 3 quadrangle,
// 4 tetrahedron, 5 hexahedron, 6 prism, 7 pyramid, 15 point.
bool elem_type(int t, ElemType& e) {
    switch (t) {
        case 15: e = {0, 1}; return true;
        case 1: e = {1, 2}; return true;
        case 2: e = {2, 3}; return true;
        case 3: e = {2, 4}; return true;
        case 4: e = {3, 4}; return true;
        case 5: e = {3, 8}; return true;
        case 6: e = {3, 6}; return true;
        case 7: e = {3, 5}; return true;
        default: return false;
    }
}

// Local faces of a cell (node positions in the Gmsh node order of the element); -1 ends a triangle.
const std::vector<std::array<int, 4>>& cell_faces(int type) {
    static const std::vector<std::array<int, 4>> tri = {{0, 1, -1, -1}, {1, 2, -1, -1}, {2, 0, -1, -1}};
    static const std::vector<std::array<int, 4>> quad = {{0, 1, -1, -1}, {1, 2, -1, -1}, {2, 3, -1, -1}, {3, 0, -1, -1}};
    static const std::vector<std::array<int, 4>> tet = {{0, 2, 1, -1}, {0, 1, 3, -1}, {0, 3, 2, -1}, {1, 2, 3, -1}};
    static const std::vector<std::array<int, 4>> hex = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3},
                                                        {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
    static const std::vector<std::array<int, 4>> prism = {{0, 2, 1, -1}, {3, 4, 5, -1}, {0, 1, 4, 3},
                                                          {1, 2, 5, 4}, {0, 3, 5, 2}};
    static const std::vector<std::array<int, 4>> pyr = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1},
                                                        {2, 3, 4, -1}, {3, 0, 4, -1}};
    switch (type) {
        case 2: return tri;
        case 3: return quad;
        case 4: return tet;
        case 5: return hex;
        case 6: return prism;
        default: return pyr;
    }
}

using Key = std::array<int64_t, 4>;  // sorted node indices, -1 padded
Key face_key(const int64_t* n, int k) {
    Key key{-1, -1, -1, -1};
    for (int i = 0; i < k; i++) key[i] = n[i];
    std::sort(key.begin(), key.begin() + k);
    return key;
}

// Whitespace tokenizer over one section of the file.
struct Tok {
    std::istringstream in;
    explicit Tok(const std::string& s) : in(s) {}
    template <typename T>
    T next() {
        T v;
        if (!(in >> v)) throw std::runtime_error("abl_fv_mesh_gmsh: truncated section");
        return v;
    }
};

std::string section(const std::string& text, const std::string& name, bool required) {
    const std::string open = "$" + name, close = "$End" + name;
    size_t a = text.find(open + "\n");
    if (a == std::string::npos) a = text.find(open + "\r\n");
    if (a == std::string::npos) {
        if (required) throw std::runtime_error("abl_fv_mesh_gmsh: missing $" + name + " section");
        return std::string();
    }
    a += open.size();
    size_t b = text.find(close, a);
    if (b == std::string::npos) throw std::runtime_error("abl_fv_mesh_gmsh: unterminated $" + name + " section");
    return text.substr(a, b - a);
}

}  // namespace
}  // namespace abl

using abl::HostMesh;

extern "C" int abl_fv_mesh_gmsh(const char* path, int32_t ordering, abl_fv_mesh_handle* out) {
    using namespace abl;
    try {
        if (!path || !out) throw std::invalid_argument("abl_fv_mesh_gmsh: null argument");
        if (ordering != ABL_FV_ORDER_NATURAL && ordering != ABL_FV_ORDER_RCM && ordering != ABL_FV_ORDER_MORTON)
            throw std::invalid_argument("abl_fv_mesh_gmsh: ordering must be natural, RCM or Morton");
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error(std::string("abl_fv_mesh_gmsh: cannot open ") + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        {
            Tok t(section(text, "MeshFormat", true));
            const std::string version = t.next<std::string>();
            const int file_type = t.next<int>();
            if (version.rfind("4.1", 0) != 0 || file_type != 0)
                throw std::runtime_error("abl_fv_mesh_gmsh: only MSH 4.1 ASCII files are supported (got version " +
                                         version + (file_type ? ", binary)" : ")"));
        }
        // entities: (dim, tag) -> first physical tag (-1 = none)
        std::map<std::pair<int, int>, int> phys;
        {
            Tok t(section(text, "Entities", true));
            int64_t cnt[4];
            for (int d = 0; d < 4; d++) cnt[d] = t.next<int64_t>();
            for (int d = 0; d < 4; d++)
                for (int64_t i = 0; i < cnt[d]; i++) {
                    const int tag = t.next<int>();
                    for (int k = 0; k < (d == 0 ? 3 : 6); k++) t.next<double>();
                    const int np = t.next<int>();
                    int first = -1;
                    for (int k = 0; k < np; k++) {
                        const int p = t.next<int>();
                        if (k == 0) first = p;
                    }
                    if (d > 0) {
                        const int nb = t.next<int>();
                        for (int k = 0; k < nb; k++) t.next<int>();
                    }
                    phys[{d, tag}] = first;
                }
        }
        // nodes: tag -> dense index
        std::vector<double> X;
        std::map<int64_t, int64_t> node_index;
        {
            Tok t(section(text, "Nodes", true));
            const int64_t nblocks = t.next<int64_t>(), nnodes = t.next<int64_t>();
            t.next<int64_t>();
            t.next<int64_t>();
            X.reserve(3 * nnodes);
            for (int64_t b = 0; b < nblocks; b++) {
                const int edim = t.next<int>();
                t.next<int>();
                const int parametric = t.next<int>();
                const int64_t n = t.next<int64_t>();
                std::vector<int64_t> tags(n);
                for (auto& g : tags) g = t.next<int64_t>();
                for (int64_t i = 0; i < n; i++) {
                    node_index[tags[i]] = (int64_t)X.size() / 3;
                    for (int d = 0; d < 3; d++) X.push_back(t.next<double>());
                    if (parametric)
                        for (int d = 0; d < edim; d++) t.next<double>();
                }
            }
        }
        // elements
        struct Elem {
            int type, dim, phys;
            std::vector<int64_t> nodes;
        };
        std::vector<Elem> elems;
        int top = 0;
        {
            Tok t(section(text, "Elements", true));
            const int64_t nblocks = t.next<int64_t>();
            t.next<int64_t>();
            t.next<int64_t>();
            t.next<int64_t>();
            for (int64_t b = 0; b < nblocks; b++) {
                const int edim = t.next<int>(), etag = t.next<int>(), etype = t.next<int>();
                const int64_t n = t.next<int64_t>();
                ElemType et;
                if (!elem_type(etype, et))
                    throw std::runtime_error("abl_fv_mesh_gmsh: unsupported element type " + std::to_string(etype) +
                                             " (first-order elements only)");
                auto it = phys.find({edim, etag});
                const int ph = it == phys.end() ? -1 : it->second;
                for (int64_t i = 0; i < n; i++) {
                    t.next<int64_t>();
                    Elem e{etype, et.dim, ph, std::vector<int64_t>(et.nodes)};
                    for (auto& v : e.nodes) {
                        auto ni = node_index.find(t.next<int64_t>());
                        if (ni == node_index.end()) throw std::runtime_error("abl_fv_mesh_gmsh: element names an unknown node");
                        v = ni->second;
                    }
                    top = std::max(top, et.dim);
                    elems.push_back(std::move(e));
                }
            }
        }
        if (top < 2) throw std::runtime_error("abl_fv_mesh_gmsh: no 2-D or 3-D cells in the file");
        const int dim = top;
        std::vector<int64_t> cell_elem;
        std::map<Key, int> bnd_phys;  // boundary element (dim - 1) -> physical tag
        for (size_t i = 0; i < elems.size(); i++) {
            if (elems[i].dim == dim) cell_elem.push_back((int64_t)i);
            else if (elems[i].dim == dim - 1 && elems[i].phys >= 0)
                bnd_phys[face_key(elems[i].nodes.data(), (int)elems[i].nodes.size())] = elems[i].phys;
        }
        const int64_t NI = (int64_t)cell_elem.size();
        if (NI >= (int64_t)1 << 31) throw std::runtime_error("abl_fv_mesh_gmsh: too many cells");
        // faces by node key
        struct FaceRec {
            Key key;
            int32_t cell, local;
        };
        std::vector<FaceRec> recs;
        for (int64_t c = 0; c < NI; c++) {
            const Elem& e = elems[cell_elem[c]];
            const auto& lf = cell_faces(e.type);
            for (int l = 0; l < (int)lf.size(); l++) {
                int64_t n[4];
                int k = 0;
                for (int q = 0; q < 4 && lf[l][q] >= 0; q++) n[k++] = e.nodes[lf[l][q]];
                recs.push_back({face_key(n, k), (int32_t)c, l});
            }
        }
        std::sort(recs.begin(), recs.end(), [](const FaceRec& a, const FaceRec& b) {
            return a.key != b.key ? a.key < b.key : a.cell < b.cell;
        });
        struct FacePair {
            int32_t a, b, la;  // b < 0: boundary; la = local face of a (its node list gives the geometry)
            Key key;
        };
        std::vector<FacePair> fp;
        for (size_t r = 0; r < recs.size();) {
            if (r + 1 < recs.size() && recs[r + 1].key == recs[r].key) {
                if (r + 2 < recs.size() && recs[r + 2].key == recs[r].key)
                    throw std::runtime_error("abl_fv_mesh_gmsh: a face is shared by more than two cells");
                fp.push_back({recs[r].cell, recs[r + 1].cell, recs[r].local, recs[r].key});
                r += 2;
            } else {
                fp.push_back({recs[r].cell, -1, recs[r].local, recs[r].key});
                r += 1;
            }
        }
        // cell geometry (file order)
        std::vector<double> vol(NI), cent(3 * NI, 0.0);
        std::vector<int64_t> source(NI);
        for (int64_t c = 0; c < NI; c++) {
            const Elem& e = elems[cell_elem[c]];
            double* xc = &cent[3 * c];
            source[c] = c;
            if (dim == 2) {
                std::vector<const double*> v;
                for (auto n : e.nodes) v.push_back(&X[3 * n]);
                polygon_geometry(v, vol[c], xc);
            } else {
                // tetrahedral fan about the vertex mean p over the triangulated faces (outward by construction
                // of the signed volumes' sum)
                double p[3] = {0, 0, 0};
                for (auto n : e.nodes)
                    for (int d = 0; d < 3; d++) p[d] += X[3 * n + d] / (double)e.nodes.size();
                double V = 0.0, cs[3] = {0, 0, 0};
                for (const auto& lf : cell_faces(e.type)) {
                    int k2 = 0;
                    const double* v[4];
                    for (int q = 0; q < 4 && lf[q] >= 0; q++) v[k2++] = &X[3 * e.nodes[lf[q]]];
                    for (int i = 1; i + 1 < k2; i++) {
                        double a[3], b[3], cc[3];
                        for (int d = 0; d < 3; d++) {
                            a[d] = v[0][d] - p[d];
                            b[d] = v[i][d] - p[d];
                            cc[d] = v[i + 1][d] - p[d];
                        }
                        const double t = (a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) +
                                          a[2] * (b[0] * cc[1] - b[1] * cc[0])) / 6.0;
                        V += t;
                        for (int d = 0; d < 3; d++) cs[d] += t * (p[d] + v[0][d] + v[i][d] + v[i + 1][d]) / 4.0;
                    }
                }
                vol[c] = std::fabs(V);
                for (int d = 0; d < 3; d++) xc[d] = cs[d] / V;
            }
            if (!(vol[c] > 0.0)) throw std::runtime_error("abl_fv_mesh_gmsh: degenerate cell");
        }
        // faces: the node loop of the left cell's local face; boundary faces take their physical tag
        PolyFaces pf;
        for (const auto& p : fp) {
            const Elem& e = elems[cell_elem[p.a]];
            const auto& lf = cell_faces(e.type)[p.la];
            int64_t nodes[4];
            int k = 0;
            for (int q = 0; q < 4 && lf[q] >= 0; q++) nodes[k++] = e.nodes[lf[q]];
            int32_t slot = -1;
            if (p.b < 0) {
                auto it = bnd_phys.find(p.key);
                if (it == bnd_phys.end()) throw std::runtime_error("abl_fv_mesh_gmsh: a boundary face has no physical group");
                slot = it->second;
            }
            pf.add(p.a, p.b, slot, nodes, k);
        }
        auto* mesh = new abl_fv_mesh();
        try {
            assemble_polytope_mesh(dim, X, vol, cent, pf, ordering, source, "abl_fv_mesh_gmsh", mesh->m);
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
