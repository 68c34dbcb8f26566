// Internal declarations shared by the host-side translation units of libablate_fv.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ablate_fv.h"

namespace abl {

// Host-side generated mesh (arrays follow abl_fv_mesh_desc).
struct HostMesh {
    int32_t dim = 3;
    int64_t n_cells = 0, n_interior = 0, n_faces = 0;
    std::vector<int32_t> face_cells;   // [F][2]
    std::vector<double> face_normal;   // [F][3]
    std::vector<double> face_centroid; // [F][3]
    std::vector<double> face_shift;    // [F][3]
    std::vector<double> cell_centroid; // [N][3]
    std::vector<double> cell_volume;   // [N]
    std::vector<int32_t> face_bc;      // [F]
    std::vector<int64_t> source;       // [NI] index of each interior cell in the producer's own numbering
    bool has_shift = false;
    bool has_bc = false;
};

// Ghost cell `ghost` of boundary face f: the centroid of `interior` reflected across the face plane and its volume
// (the DMPlexComputeGeometryFVM ghost-cell convention).
void finish_ghost(HostMesh& m, int64_t f, int32_t interior, int32_t ghost);

// Sets the calling thread's global error message (returned by abl_fv_last_error(NULL)).
void set_global_error(const std::string& msg);

// Morton (Z-order) order of n points xc[3 n] (21 bits per axis over their bounding cube); returns perm[new] = old.
std::vector<int64_t> morton_order(const std::vector<double>& xc, int64_t n);

// Reverse Cuthill-McKee order of a cell graph given as CSR adjacency; returns perm[new] = old.
std::vector<int64_t> rcm_order(int64_t n, const std::vector<int64_t>& off, const std::vector<int32_t>& adj);

// Faces of a mesh given as polytopes, before numbering: left and right input cell (right < 0 on the boundary),
// the boundary-condition slot of boundary faces (-1 elsewhere) and each face's vertex loop (2 vertices in 2-D).
struct PolyFaces {
    std::vector<int32_t> a, b, bc;
    std::vector<int64_t> off{0}, vtx;
    void add(int32_t left, int32_t right, int32_t slot, const int64_t* v, int nv) {
        a.push_back(left);
        b.push_back(right);
        bc.push_back(slot);
        vtx.insert(vtx.end(), v, v + nv);
        off.push_back((int64_t)vtx.size());
    }
};

// Polygon (2-D) area and centroid (shoelace about the first vertex of the loop; c[2] = 0).
void polygon_geometry(const std::vector<const double*>& v, double& area, double* c);

// Planar polygon face in 3-D: Newell area vector n (orientation of the loop) and the fan centroid c.
void face_geometry3(const std::vector<const double*>& v, double* n, double* c);

// Builds the finite-volume mesh m from polytopes: interior cells numbered by `ordering` (natural / RCM / Morton
// on the centroids), faces numbered by their left cell (the lower cell number, boundary faces after a cell's
// interior ones) with normals of length |face| pointing from left to right, one ghost cell per boundary face
// (finish_ghost).  X holds vertex coordinates [nv][3]; vol / cent [NI] / [NI][3] the interior cells' volumes and
// centroids in input order; source[c] the producer's index of input cell c.  Errors name `who`.
void assemble_polytope_mesh(int dim, const std::vector<double>& X, const std::vector<double>& vol,
                            const std::vector<double>& cent, const PolyFaces& pf, int ordering,
                            const std::vector<int64_t>& source, const char* who, HostMesh& m);

// Minimum-norm least-squares weights: w[i*dim+d] = pinv(dx)[d][i] for the n x dim matrix dx.
void lsq_pinv_weights(int n, int dim, const double* dx, double* w);

}  // namespace abl

struct abl_fv_mesh {
    abl::HostMesh m;
};
