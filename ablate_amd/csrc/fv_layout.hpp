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

// Cell tiles for the fused face-flux + gather kernel: tile t owns interior cells [t*T, min((t+1)*T, NI)).
//   start[t] .. start[t+1]   its faces (every face with a cell in the tile), ascending face number
//   slot[k * NI + c]         (tile-local index of the k-th face of c) << 1 | side, -1 padding; KM slots
//   nf_max                   largest face count of a tile (sizes the workgroup's LDS)
struct Tiles {
    int T = 0;
    int64_t n = 0;
    int nf_max = 0;
    std::vector<int32_t> start, faces, slot;
};
void build_tiles(const Layout& L, int KM, int T, Tiles& out);

// Cell tiles of the fused strategy (gradients + face fluxes + gather in one kernel): tile t owns interior cells
// [t*T, min((t+1)*T, NI)); its halo is the interior cells outside the tile that share a face with it.
//   fstart / ffaces   the tile's faces (every face with a cell in the tile), ascending face number
//   fcode             per tile face: left side code | right side code << 16; a code is a tile cell's local index,
//                     T + the halo face entry of that side, or 0xFFFF for a ghost cell
//   fslot[k * NI + c] (tile-local index of the k-th face of c) << 1 | side, -1 padding; KM slots
//   hstart / hcell    the tile's halo cells, ascending
//   hinfo             per halo cell: bit k set when its k-th face is a tile face, | its first face entry << 8 (one
//                     entry per such face, in slot order)
//   nfm / nem         largest face / halo face-entry count of a tile (size the workgroup's LDS)
struct FusedTiles {
    int T = 0, nfm = 0, nem = 0;
    int64_t n = 0;
    std::vector<int32_t> fstart, ffaces, fcode, fslot, hstart, hcell, hinfo;
};
void build_fused_tiles(const Layout& L, int KM, int T, FusedTiles& out);
// Checks every index the fused kernel takes from the lists against the mesh and the list bounds; throws
// std::logic_error naming the first violated invariant.
void check_fused_tiles(const Layout& L, int KM, const FusedTiles& ft);

// Plan of the marching strategy: tiles of T consecutive interior cells (tile t owns [t*T, min((t+1)*T, NI))) are
// walked in sequences, one workgroup per sequence.  Each step of a sequence handles one tile ("cur") and computes
// the cell records of the tile the next step handles ("next"), so faces between consecutive tiles of a sequence are
// evaluated once and no record is computed twice inside a sequence; face neighbours in other tiles ("halo") are
// recomputed per step as in the fused strategy.  A sequence starts with a prologue: the records of its first tile
// and of a "prologue tile" sharing faces with it (the tile before it, or its periodic partner), and the fluxes of
// the faces between the two.
//   seq[w] .. seq[w+1]          steps of sequence w; pro[w] its prologue tile or -1
//   stile[s], snext[s]          tile handled by step s, tile whose records step s computes (-1: none)
//   fstart / ffaces / fcode     per step: faces evaluated (ascending), side codes lc | rc << 16 with
//                               [0, T) a cur cell, [T, 2T) a next cell, [2T, 2T + nem) a halo entry, 0xFFFF a ghost
//   pfstart / pffaces / pfcode  per sequence: prologue faces, codes [0, T) first-tile cell, [T, 2T) prologue cell
//   hstart / hcell / hinfo      per step: halo cells and entries (slot mask | first entry << 8), as FusedTiles
//   gslot[k * NI + c]           gather code of slot k of cell c: list index << 2 | 2 (if evaluated by the previous
//                               step or the prologue) | side, -1 padding
struct MarchPlan {
    int T = 0, nfm = 0, nem = 0;
    int64_t n_seq = 0, n_steps = 0;
    std::vector<int32_t> seq, pro, stile, snext, fstart, ffaces, fcode, pfstart, pffaces, pfcode, hstart, hcell, hinfo,
        gslot;
};
// max_steps bounds the steps of one sequence (chains of face-sharing consecutive tiles are split evenly).
void build_march_plan(const Layout& L, int KM, int T, int max_steps, MarchPlan& out);
// Checks every index the marching kernel takes from the plan, and that every gather code names the face of its slot
// in the list (and buffer) that holds it; throws std::logic_error naming the first violated invariant.
void check_march_plan(const Layout& L, int KM, const MarchPlan& p);

// True when every interior cell's least-squares stencil (its interior face neighbours, offsets from fdx) has full
// rank, so that the normal equations the fused kernel solves give the pseudo-inverse solution.
bool lsq_stencils_full_rank(const Layout& L);

}  // namespace abl
