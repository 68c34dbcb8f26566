// Strategy 4 (fused): gradients, limiter, viscous terms, face fluxes and the residual gather of one cell tile in
// one kernel, with no cell record and no face flux in HBM.  Included by fv_solver.hip after the face-flux helpers
// (SideData, load_side, ghost_visc, face_flux, aux_of), inside namespace abl.
//
// One workgroup per tile of T consecutive interior cells (8x4x4 blocks of a brick-ordered box).  Per tile:
//   phase 1  every cell of the tile and of its halo (interior face neighbours outside the tile) computes its
//            least-squares gradients (CellInterpolant::ComputeFieldGradients, cellInterpolant.cpp:330-413), their
//            Barth-Jespersen limiter (:415-487) and the unlimited aux gradients of the face-interpolant viscous flux
//            (stencils/leastSquaresAverage.cpp:5-155).  Tile cells keep their record (state, limited gradients,
//            tau / k grad T) in LDS; a halo cell stores, for each of its faces in the tile, what the face needs from
//            its side: the state projected to the face (ProjectToFace, cellInterpolant.cpp:890-950), its velocity
//            and its viscous terms.
//   phase 2  every face of the tile is evaluated once from LDS into LDS (faces on the tile border by both tiles).
//   phase 3  each tile cell gathers its faces in slot order: -F/V as left cell, +F/V as right cell
//            (cellInterpolant.cpp:604-629), and writes its residual.
// The least-squares weights are not read from HBM: each cell solves its normal equations
// (sum_k dx_k dx_k^T) g = sum_k dx_k (u_k - u_c) from the offsets dx_k = x_nbr - x_c, which the per-face
// offsets x_face - x_cell of both sides give (the same minimum-norm solution as the weights' pseudo-inverse for a
// full-rank stencil; the strategy is offered only when every stencil has full rank).  The face offsets are also the
// limiter's projection vectors and the reconstruction's, so the geometry a tile reads is its faces' records.
#pragma once

// (included inside namespace abl)

// Per-tile lists of the fused strategy (built on the host by build_fused_tiles).
struct FusedView {
    const int32_t* __restrict__ fstart;  // [n + 1]: faces of tile t at [fstart[t], fstart[t + 1])
    const int32_t* __restrict__ ffaces;  // face numbers, ascending within a tile
    const int32_t* __restrict__ fcode;   // per tile face: left side code | right side code << 16
    const int32_t* __restrict__ fslot;   // [KM][NI]: tile-local face index << 1 | side, -1 padding
    const int32_t* __restrict__ hstart;  // [n + 1]: halo cells of tile t at [hstart[t], hstart[t + 1])
    const int32_t* __restrict__ hcell;   // halo cell numbers
    const int32_t* __restrict__ hinfo;   // per halo cell: mask of its slots whose face is a tile face (bits 0-7)
                                         // | index of its first face entry << 8 (entries follow in slot order)
    int T;    // tile cells
    int nfm;  // largest face count of a tile
    int nem;  // largest halo face-entry count of a tile
};
// Side codes: [0, T) a tile cell, [T, T + nem) a halo face entry, kFusedGhost a ghost cell (read from U).
constexpr int kFusedGhost = 0xFFFF;

template <int DIM, bool RECON, bool VISC>
struct FusedLayout {
    static constexpr int NE = DIM + 2;
    static constexpr int NT = ViscTerms<DIM>::NT;
    static constexpr int G0 = NE;                               // limited gradients (p * DIM + d)
    static constexpr int T0 = G0 + (RECON ? NE * DIM : 0);      // viscous terms
    static constexpr int RW = T0 + (VISC ? NT : 0);             // words of a tile cell's record
    static constexpr int HV = NE;                               // halo entry: projected state [0, NE)
    static constexpr int HT = NE + (VISC ? DIM : 0);            // velocity [NE, HT), viscous terms [HT, HW)
    static constexpr int HW = HT + (VISC ? NT : 0);             // words of a halo face entry
};

// Solves the symmetric normal equations M x = b for each right-hand side (M given by its upper triangle, packed
// row by row): x = adj(M) b / det(M).
template <int DIM>
struct SymSolve {
    double a[DIM * (DIM + 1) / 2];  // adjugate, packed like M
    double idet;
    __device__ inline void init(const double* M) {
        if constexpr (DIM == 3) {
            const double m00 = M[0], m01 = M[1], m02 = M[2], m11 = M[3], m12 = M[4], m22 = M[5];
            a[0] = m11 * m22 - m12 * m12;
            a[1] = m02 * m12 - m01 * m22;
            a[2] = m01 * m12 - m02 * m11;
            a[3] = m00 * m22 - m02 * m02;
            a[4] = m01 * m02 - m00 * m12;
            a[5] = m00 * m11 - m01 * m01;
            idet = 1.0 / (m00 * a[0] + m01 * a[1] + m02 * a[2]);
        } else {
            a[0] = M[2];
            a[1] = -M[1];
            a[2] = M[0];
            idet = 1.0 / (M[0] * M[2] - M[1] * M[1]);
        }
    }
    __device__ inline void apply(const double* b, double* x) const {
        if constexpr (DIM == 3) {
            x[0] = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * idet;
            x[1] = (a[1] * b[0] + a[3] * b[1] + a[4] * b[2]) * idet;
            x[2] = (a[2] * b[0] + a[4] * b[1] + a[5] * b[2]) * idet;
        } else {
            x[0] = (a[0] * b[0] + a[1] * b[1]) * idet;
            x[1] = (a[1] * b[0] + a[2] * b[1]) * idet;
        }
    }
};
template <int DIM>
__device__ inline void sym_add_outer(double* M, const double* v) {
    int i = 0;
#pragma unroll
    for (int r = 0; r < DIM; r++)
#pragma unroll
        for (int c = r; c < DIM; c++) M[i++] += v[r] * v[c];
}

// Phase 1 for one cell: state uc, limited gradients g, viscous terms tq and each slot's face offset x_f - x_c.
// Slot order, ghost handling and the limiter follow gradients_body (the two-pass strategies' gradient kernel).
template <int DIM, bool RECON, bool VISC, int LIM, int KM>
__device__ __forceinline__ void fused_cell(const MeshView<double>& m, const Phys& ph, const double* __restrict__ U,
                                           int64_t c, double* uc, double (*g)[DIM], double* tq, double (*xf)[DIM]) {
    constexpr int NE = DIM + 2;
    constexpr int NA = DIM + 1;
    constexpr int NM = DIM * (DIM + 1) / 2;
    const int64_t N = m.N, NI = m.NI;
    const bool aux_ghosts = m.fwa != nullptr;  // the face-stencil (aux) set includes ghost neighbours
    double ac[NA];
#pragma unroll
    for (int p = 0; p < NE; p++) uc[p] = U[p * N + c];
    if (VISC) aux_of<DIM>(uc, ph, ac);
    double M[NM], Mg[NM], b[NE][DIM], ba[NA][DIM], umin[NE], umax[NE];
    bool valid[KM];
#pragma unroll
    for (int i = 0; i < NM; i++) {
        M[i] = 0.0;
        Mg[i] = 0.0;
    }
#pragma unroll
    for (int p = 0; p < NE; p++) {
        umin[p] = uc[p];
        umax[p] = uc[p];
#pragma unroll
        for (int d = 0; d < DIM; d++) b[p][d] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < NA; q++)
#pragma unroll
        for (int d = 0; d < DIM; d++) ba[q][d] = 0.0;

#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = k < m.K ? m.ell[k * NI + c] : -1;
        const int32_t nb = m.snbr[k * NI + c];
        valid[k] = e >= 0;
        const int64_t f = valid[k] ? (int64_t)(e >> 1) : 0;
        const int s = e & 1;
        const int64_t n = valid[k] ? (int64_t)nb : c;
        double x[2][DIM];
        load_face_pair<DIM>(m.fdx, f, x);
        double dxn[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            xf[k][d] = s ? x[1][d] : x[0][d];
            dxn[d] = s ? x[1][d] - x[0][d] : x[0][d] - x[1][d];  // x_n - x_c = (x_f - x_c) - (x_f - x_n)
        }
        double un[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) un[p] = U[p * N + n];
        if (!valid[k]) continue;
        const bool bnd = n >= NI;
        if (RECON) {
            if (LIM == kLimBJ) {
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    umin[p] = fmin(umin[p], un[p]);
                    umax[p] = fmax(umax[p], un[p]);
                }
            }
        }
        if (!bnd) {
            sym_add_outer<DIM>(M, dxn);
            if (RECON)
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    const double delta = un[p] - uc[p];
#pragma unroll
                    for (int d = 0; d < DIM; d++) b[p][d] += dxn[d] * delta;
                }
        } else if (VISC && aux_ghosts) {
            sym_add_outer<DIM>(Mg, dxn);
        }
        if (VISC && (!bnd || aux_ghosts)) {
            double an[NA];
            aux_of<DIM>(un, ph, an);
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const double delta = an[q] - ac[q];
#pragma unroll
                for (int d = 0; d < DIM; d++) ba[q][d] += dxn[d] * delta;
            }
        }
    }
    SymSolve<DIM> sc;
    sc.init(M);
    if (RECON) {
#pragma unroll
        for (int p = 0; p < NE; p++) sc.apply(b[p], g[p]);
        if (LIM == kLimBJ) {
            double dpos[NE], dneg[NE];
#pragma unroll
            for (int p = 0; p < NE; p++) {
                dpos[p] = 0.0;
                dneg[p] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < KM; k++) {
                if (!valid[k]) continue;
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    double D = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) D += g[p][d] * xf[k][d];
                    dpos[p] = fmax(dpos[p], D);
                    dneg[p] = fmin(dneg[p], D);
                }
            }
            double phi[NE];
#pragma unroll
            for (int p = 0; p < NE; p++) {
                double v = 1.0;
                if (dpos[p] > 0.0) v = fmin(v, (umax[p] - uc[p]) / dpos[p]);
                if (dneg[p] < 0.0) v = fmin(v, (umin[p] - uc[p]) / dneg[p]);
                phi[p] = v;
            }
            bool cancel = false;
#pragma unroll
            for (int p = 0; p < NE; p++)
                if (phi[p] == 0.0)
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        if (fabs(g[p][d]) > ph.max_lim_grad) cancel = true;
#pragma unroll
            for (int p = 0; p < NE; p++)
#pragma unroll
                for (int d = 0; d < DIM; d++) g[p][d] *= cancel ? 0.0 : phi[p];
        }
    }
    if (VISC) {
        double ga[NA][DIM];
        if (aux_ghosts) {
            double Ma[NM];
#pragma unroll
            for (int i = 0; i < NM; i++) Ma[i] = M[i] + Mg[i];
            SymSolve<DIM> sa;
            sa.init(Ma);
#pragma unroll
            for (int q = 0; q < NA; q++) sa.apply(ba[q], ga[q]);
        } else {
#pragma unroll
            for (int q = 0; q < NA; q++) sc.apply(ba[q], ga[q]);
        }
        stress_heat<double, DIM>(ph.mu, ph.k, ga, tq);
    }
}

// Side data of a face from LDS: a tile cell's record (projected with the face offset dx) or a halo face entry.
template <int DIM, bool RECON, bool VISC>
__device__ inline void fused_side(const double* rec, int T, const double* hal, int nem, int code, const double* dx,
                                  SideData<DIM, VISC>& sd) {
    using FL = FusedLayout<DIM, RECON, VISC>;
    constexpr int NE = DIM + 2;
    if (code < T) {
        double r[FL::RW];
#pragma unroll
        for (int w = 0; w < FL::RW; w++) r[w] = rec[w * T + code];
#pragma unroll
        for (int p = 0; p < NE; p++) {
            double a = 0.0;
            if (RECON)
#pragma unroll
                for (int d = 0; d < DIM; d++) a += r[FL::G0 + p * DIM + d] * dx[d];
            sd.uf[p] = r[p] + a;
        }
        if (VISC)
#pragma unroll
            for (int q = 0; q < FL::NT; q++) sd.tq[q] = r[FL::T0 + q];
        side_cell_values<DIM, VISC>(r, sd);
    } else {
        const int e = code - T;
#pragma unroll
        for (int p = 0; p < NE; p++) sd.uf[p] = hal[p * nem + e];
        sd.rho = 0.0;  // only the species diffusion coefficient reads it
        sd.T = 0.0;    // only ghost_visc reads it, and a halo cell is never next to a ghost face of the tile
        if (VISC) {
#pragma unroll
            for (int d = 0; d < DIM; d++) sd.vel[d] = hal[(FL::HV + d) * nem + e];
#pragma unroll
            for (int q = 0; q < FL::NT; q++) sd.tq[q] = hal[(FL::HT + q) * nem + e];
        }
    }
}

template <int DIM, bool RECON, bool VISC, int LIM, int KM>
__global__ __launch_bounds__(256) void k_fused(MeshView<double> m, Phys ph, const double* __restrict__ U, FusedView fv,
                                               double* __restrict__ R, int accumulate) {
    using FL = FusedLayout<DIM, RECON, VISC>;
    using SD = SideData<DIM, VISC>;
    constexpr int NE = DIM + 2;
    extern __shared__ double lds[];
    const int T = fv.T, nem = fv.nem, nfm = fv.nfm;
    double* rec = lds;                    // [RW][T]
    double* hal = rec + FL::RW * T;       // [HW][nem]
    double* flx = hal + FL::HW * nem;     // [NE][nfm]
    const int64_t t = xcd_block_index();
    const int64_t c0 = t * T;
    const int64_t N = m.N, NI = m.NI, F = m.F;
    const int nown = (int)(NI - c0 < (int64_t)T ? NI - c0 : (int64_t)T);
    const int32_t h0 = fv.hstart[t], nh = fv.hstart[t + 1] - h0;

    // phase 1: records of the tile cells, face entries of the halo cells
    for (int i = threadIdx.x; i < nown + nh; i += blockDim.x) {
        const bool own = i < nown;
        const int64_t c = own ? c0 + i : (int64_t)fv.hcell[h0 + i - nown];
        double uc[NE], g[NE][DIM], tq[ViscTerms<DIM>::NT], xf[KM][DIM];
        fused_cell<DIM, RECON, VISC, LIM, KM>(m, ph, U, c, uc, g, tq, xf);
        if (own) {
#pragma unroll
            for (int p = 0; p < NE; p++) rec[p * T + i] = uc[p];
            if (RECON)
#pragma unroll
                for (int p = 0; p < NE; p++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) rec[(FL::G0 + p * DIM + d) * T + i] = g[p][d];
            if (VISC)
#pragma unroll
                for (int q = 0; q < FL::NT; q++) rec[(FL::T0 + q) * T + i] = tq[q];
        } else {
            const int32_t info = fv.hinfo[h0 + i - nown];
            int e = info >> 8;
            double vel[DIM];
            if (VISC) {
                const double irho = 1.0 / uc[kRho];
#pragma unroll
                for (int d = 0; d < DIM; d++) vel[d] = uc[kRhoU + d] * irho;
            }
#pragma unroll
            for (int k = 0; k < KM; k++) {
                if (!((info >> k) & 1)) continue;
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    double a = 0.0;
                    if (RECON)
#pragma unroll
                        for (int d = 0; d < DIM; d++) a += g[p][d] * xf[k][d];
                    hal[p * nem + e] = uc[p] + a;
                }
                if (VISC) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) hal[(FL::HV + d) * nem + e] = vel[d];
#pragma unroll
                    for (int q = 0; q < FL::NT; q++) hal[(FL::HT + q) * nem + e] = tq[q];
                }
                e++;
            }
        }
    }
    __syncthreads();

    // phase 2: the tile's faces
    const int32_t f0 = fv.fstart[t], nf = fv.fstart[t + 1] - f0;
    for (int j = threadIdx.x; j < nf; j += blockDim.x) {
        const int64_t f = fv.ffaces[f0 + j];
        const int32_t code = fv.fcode[f0 + j];
        const int lc = code & 0xFFFF, rc = (int)((uint32_t)code >> 16);
        double nh[3], area, x[2][DIM];
        load_face_na(m.fna, f, nh, area);
        load_face_pair<DIM>(m.fdx, f, x);
        SD sl, sr;
        fused_side<DIM, RECON, VISC>(rec, T, hal, nem, lc, x[0], sl);
        if (rc == kFusedGhost) {
            const int64_t Rc = m.fcell[F + f];
            load_side<DIM, RECON, VISC>(U, nullptr, N, Rc, true, x[1], sr);
            if (VISC) {
                double w[2][DIM];
                load_face_pair<DIM>(m.fwa, f, w);
                ghost_visc<DIM, VISC>(ph, w[1], sl, sr);
            }
        } else {
            fused_side<DIM, RECON, VISC>(rec, T, hal, nem, rc, x[1], sr);
        }
        double flux[NE], cl, cr, cd;
        face_flux<DIM, VISC>(ph, nh, area, sl, sr, flux, cl, cr, cd);
#pragma unroll
        for (int p = 0; p < NE; p++) flx[p * nfm + j] = flux[p];
    }
    __syncthreads();

    // phase 3: each tile cell gathers its faces in slot order
    if ((int)threadIdx.x >= nown) return;
    const int64_t c = c0 + threadIdx.x;
    int li[KM];
    double sg[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = fv.fslot[k * NI + c];
        li[k] = e >= 0 ? (e >> 1) : 0;
        sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) acc += sg[k] * flx[p * nfm + li[k]];
        const double r = acc * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
}


// ================================================================================================ marching strategy
//
// One workgroup per sequence of tiles (MarchPlan).  Step s handles tile t_s ("cur") and holds two record slots in
// LDS: cur's records (computed by the previous step, or by the prologue) and the records of the tile the next step
// handles ("next"), which step s computes.  Per step:
//   A  records of the next tile into the other slot, face entries of the step's halo cells (as k_fused phase 1)
//   B  every face of cur not shared with the previous tile, from LDS, into the step's flux buffer
//   C  cur's cells gather their faces in slot order from this step's buffer or the previous step's (the faces
//      shared with the previous tile), -F/V as left cell, +F/V as right cell (cellInterpolant.cpp:604-629)
// so faces between consecutive tiles of a sequence are evaluated once and no cell's gradients are computed twice
// inside a sequence.  Two barriers per step: A -> B and B -> C; the next step's A writes only what B has finished
// reading, and its barrier orders C before the next B overwrites the previous-step buffer.

struct MarchView {
    const int32_t* __restrict__ seq;      // [n_seq + 1]
    const int32_t* __restrict__ pro;      // [n_seq]
    const int32_t* __restrict__ stile;    // [n_steps]
    const int32_t* __restrict__ snext;    // [n_steps]
    const int32_t* __restrict__ fstart;   // [n_steps + 1]
    const int32_t* __restrict__ ffaces;
    const int32_t* __restrict__ fcode;
    const int32_t* __restrict__ pfstart;  // [n_seq + 1]
    const int32_t* __restrict__ pffaces;
    const int32_t* __restrict__ pfcode;
    const int32_t* __restrict__ hstart;   // [n_steps + 1]
    const int32_t* __restrict__ hcell;
    const int32_t* __restrict__ hinfo;
    const int32_t* __restrict__ gslot;    // [KM][NI]
    int T, nfm, nem;
};
constexpr int kMarchBlock = 128;

// Record of a tile cell into a slot ([RW][T], word-major so that a wave's stores are consecutive).
template <int DIM, bool RECON, bool VISC>
__device__ inline void march_store_record(double* slot, int T, int i, const double* uc, const double (*g)[DIM],
                                          const double* tq) {
    using FL = FusedLayout<DIM, RECON, VISC>;
    constexpr int NE = DIM + 2;
#pragma unroll
    for (int p = 0; p < NE; p++) slot[p * T + i] = uc[p];
    if (RECON)
#pragma unroll
        for (int p = 0; p < NE; p++)
#pragma unroll
            for (int d = 0; d < DIM; d++) slot[(FL::G0 + p * DIM + d) * T + i] = g[p][d];
    if (VISC)
#pragma unroll
        for (int q = 0; q < FL::NT; q++) slot[(FL::T0 + q) * T + i] = tq[q];
}

// Side data of a face from LDS: code [0, T) a cell of slot `ra`, [T, 2T) of slot `rb`, [2T, ..) a halo entry.
template <int DIM, bool RECON, bool VISC>
__device__ inline void march_side(const double* ra, const double* rb, int T, const double* hal, int nem, int code,
                                  const double* dx, SideData<DIM, VISC>& sd) {
    if (code < 2 * T) {
        const bool second = code >= T;
        fused_side<DIM, RECON, VISC>(second ? rb : ra, T, hal, nem, second ? code - T : code, dx, sd);
    } else {
        fused_side<DIM, RECON, VISC>(ra, T, hal, nem, code - T, dx, sd);  // fused_side: code >= T is entry code - T
    }
}

// Face j of a list: flux from the two sides' LDS data (ghost right sides from U) into flx[p * nfm + j].
template <int DIM, bool RECON, bool VISC>
__device__ inline void march_face(const MeshView<double>& m, const Phys& ph, const double* __restrict__ U, int64_t f,
                                  int32_t code, const double* ra, const double* rb, int T, const double* hal, int nem,
                                  double* flx, int nfm, int j) {
    using SD = SideData<DIM, VISC>;
    constexpr int NE = DIM + 2;
    const int lc = code & 0xFFFF, rc = (int)((uint32_t)code >> 16);
    double nh[3], area, x[2][DIM];
    load_face_na(m.fna, f, nh, area);
    load_face_pair<DIM>(m.fdx, f, x);
    SD sl, sr;
    march_side<DIM, RECON, VISC>(ra, rb, T, hal, nem, lc, x[0], sl);
    if (rc == kFusedGhost) {
        const int64_t Rc = m.fcell[m.F + f];
        load_side<DIM, RECON, VISC>(U, nullptr, m.N, Rc, true, x[1], sr);
        if (VISC) {
            double w[2][DIM];
            load_face_pair<DIM>(m.fwa, f, w);
            ghost_visc<DIM, VISC>(ph, w[1], sl, sr);
        }
    } else {
        march_side<DIM, RECON, VISC>(ra, rb, T, hal, nem, rc, x[1], sr);
    }
    double flux[NE], cl, cr, cd;
    face_flux<DIM, VISC>(ph, nh, area, sl, sr, flux, cl, cr, cd);
#pragma unroll
    for (int p = 0; p < NE; p++) flx[p * nfm + j] = flux[p];
}

template <int DIM, bool RECON, bool VISC, int LIM, int KM>
__global__ __launch_bounds__(kMarchBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_march(
    MeshView<double> m, Phys ph, const double* __restrict__ U, MarchView mv, double* __restrict__ R, int accumulate) {
    using FL = FusedLayout<DIM, RECON, VISC>;
    constexpr int NE = DIM + 2;
    extern __shared__ double lds[];
    const int T = mv.T, nem = mv.nem, nfm = mv.nfm;
    double* rec = lds;                      // [2][RW][T]
    double* hal = rec + 2 * FL::RW * T;     // [HW][nem]
    double* flx = hal + FL::HW * nem;       // [2][NE][nfm]
    const int64_t w = xcd_block_index();
    const int64_t N = m.N, NI = m.NI;
    const int tid = threadIdx.x;
    const int32_t s0 = mv.seq[w], s1 = mv.seq[w + 1];
    auto tile_cells = [&](int64_t t) { return (int)(NI - t * T < (int64_t)T ? NI - t * T : (int64_t)T); };

    // prologue: records of the first tile (slot 0) and of the prologue tile (slot 1), then their shared faces into
    // flux buffer 1 (the buffer step 0 reads as its previous step's)
    {
        const int64_t t0 = mv.stile[s0], tp = mv.pro[w];
        const int n0 = tile_cells(t0), np = tp >= 0 ? tile_cells(tp) : 0;
        for (int i = tid; i < n0 + np; i += blockDim.x) {
            const bool first = i < n0;
            const int li = first ? i : i - n0;
            const int64_t c = (first ? t0 : tp) * T + li;
            double uc[NE], g[NE][DIM], tq[ViscTerms<DIM>::NT], xf[KM][DIM];
            fused_cell<DIM, RECON, VISC, LIM, KM>(m, ph, U, c, uc, g, tq, xf);
            march_store_record<DIM, RECON, VISC>(rec + (first ? 0 : FL::RW * T), T, li, uc, g, tq);
        }
        __syncthreads();
        const int32_t p0 = mv.pfstart[w], np_f = mv.pfstart[w + 1] - p0;
        for (int j = tid; j < np_f; j += blockDim.x)
            march_face<DIM, RECON, VISC>(m, ph, U, mv.pffaces[p0 + j], mv.pfcode[p0 + j], rec, rec + FL::RW * T, T, hal,
                                         nem, flx + NE * nfm, nfm, j);
        __syncthreads();
    }

    for (int32_t s = s0; s < s1; s++) {
        const int par = (s - s0) & 1;
        double* rc = rec + par * FL::RW * T;        // cur tile
        double* ro = rec + (par ^ 1) * FL::RW * T;  // next tile
        double* fc = flx + par * NE * nfm;          // this step's fluxes
        double* fo = flx + (par ^ 1) * NE * nfm;    // the previous step's (or the prologue's)
        // A: next tile's records, halo entries
        {
            const int64_t tn = mv.snext[s];
            const int nn = tn >= 0 ? tile_cells(tn) : 0;
            const int32_t h0 = mv.hstart[s], nh = mv.hstart[s + 1] - h0;
            for (int i = tid; i < nn + nh; i += blockDim.x) {
                const bool own = i < nn;
                const int64_t c = own ? tn * T + i : (int64_t)mv.hcell[h0 + i - nn];
                double uc[NE], g[NE][DIM], tq[ViscTerms<DIM>::NT], xf[KM][DIM];
                fused_cell<DIM, RECON, VISC, LIM, KM>(m, ph, U, c, uc, g, tq, xf);
                if (own) {
                    march_store_record<DIM, RECON, VISC>(ro, T, i, uc, g, tq);
                } else {
                    const int32_t info = mv.hinfo[h0 + i - nn];
                    int e = info >> 8;
                    double vel[DIM];
                    if (VISC) {
                        const double irho = 1.0 / uc[kRho];
#pragma unroll
                        for (int d = 0; d < DIM; d++) vel[d] = uc[kRhoU + d] * irho;
                    }
#pragma unroll
                    for (int k = 0; k < KM; k++) {
                        if (!((info >> k) & 1)) continue;
#pragma unroll
                        for (int p = 0; p < NE; p++) {
                            double a = 0.0;
                            if (RECON)
#pragma unroll
                                for (int d = 0; d < DIM; d++) a += g[p][d] * xf[k][d];
                            hal[p * nem + e] = uc[p] + a;
                        }
                        if (VISC) {
#pragma unroll
                            for (int d = 0; d < DIM; d++) hal[(FL::HV + d) * nem + e] = vel[d];
#pragma unroll
                            for (int q = 0; q < FL::NT; q++) hal[(FL::HT + q) * nem + e] = tq[q];
                        }
                        e++;
                    }
                }
            }
        }
        __syncthreads();
        // B: this step's faces
        {
            const int32_t f0 = mv.fstart[s], nf = mv.fstart[s + 1] - f0;
            for (int j = tid; j < nf; j += blockDim.x)
                march_face<DIM, RECON, VISC>(m, ph, U, mv.ffaces[f0 + j], mv.fcode[f0 + j], rc, ro, T, hal, nem, fc,
                                             nfm, j);
        }
        __syncthreads();
        // C: gather of the cur tile's cells
        {
            const int64_t tc = mv.stile[s];
            if (tid < tile_cells(tc)) {
                const int64_t c = tc * T + tid;
                int li[KM];
                bool prev[KM];
                double sg[KM];
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    const int32_t e = mv.gslot[k * NI + c];
                    li[k] = e >= 0 ? (e >> 2) : 0;
                    prev[k] = e >= 0 && ((e >> 1) & 1);
                    sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
                }
                const double iv = m.inv_vol[c];
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < KM; k++)
                        if (sg[k] != 0.0) acc += sg[k] * (prev[k] ? fo : fc)[p * nfm + li[k]];
                    const double r = acc * iv;
                    R[p * N + c] = accumulate ? R[p * N + c] + r : r;
                }
            }
        }
    }
}
