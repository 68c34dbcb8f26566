// Host-side device-layout description of a finite-volume mesh for the RHS kernels.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/ablate_fv.h"

namespace abl {

constexpr int kMaxFacesPerCell = 8;  // width limit of the cell -> face table (hex 6, quad 4, tet 4)
constexpr int kMaxBcSlots = 16;      // boundary-condition slots
constexpr int kMaxSpecies = 64;

struct BoundaryFace {
    int32_t face;
    int32_t slot;
};

struct Layout {
    int dim = 3, K = 0, ns = 0, nc = 0;
    int64_t N = 0, NI = 0, F = 0;
    bool has_boundary = false;
    std::vector<int32_t> ell;      // [K][NI]
    std::vector<int32_t> fcell;    // [2][F]
    std::vector<double> fna;       // [F][4]: unit normal, area
    std::vector<double> fdx;       // [F][2][dim]
    std::vector<double> fw;        // [F][2][dim] (empty when unused)
    std::vector<double> fwa;       // [F][2][dim] (empty: same as fw)
    std::vector<double> inv_vol;   // [N]
    std::vector<BoundaryFace> bfaces;
};

// Validates the descriptor and fills the layout; throws std::invalid_argument on bad input.
void build_layout(const abl_fv_mesh_desc& mesh, const abl_fv_physics_desc& phys, Layout& out);

}  // namespace abl
