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

// Minimum-norm least-squares weights: w[i*dim+d] = pinv(dx)[d][i] for the n x dim matrix dx.
void lsq_pinv_weights(int n, int dim, const double* dx, double* w);

}  // namespace abl

struct abl_fv_mesh {
    abl::HostMesh m;
};
